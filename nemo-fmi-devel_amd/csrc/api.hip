// api.hip -- the extern "C" boundary of include/nemo_hip.h.
//
// Library state: one HIP stream, the registered domain (masks as per-column
// level intervals, k-indices, FP64 and FP32 copies of the metrics), library
// owned scratch arrays, the per-routine timers and, in host-pointer mode, a
// pool of staging buffers.  Every entry point validates its arguments,
// stages host arrays if needed, launches the kernels of tra.hip / dyn.hip /
// spg.hip and returns a status; no exception or abort crosses the ABI.
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "common.hpp"
#include "kernels.hpp"

namespace nh {
static char g_err[NH_MAXSTR] = "";
static bool g_init = false;
static int g_device = -1;
static hipStream_t g_stream = nullptr;
static int g_ptr_mode = NEMO_HIP_PTR_HOST;
// a context: one registered domain with its scratch and the state the
// reference keeps in module variables between calls; context 0 exists from
// init, others come from nemo_hip_context_create (ensembles of domains)
struct Ctx {
    bool have_domain = false;
    DevDomain dom{};
    nemo_hip_domain_t hdom{};
    std::vector<void *> dom_allocs;
    void *scratch3[S_NSCRATCH3] = {};
    double *scratch2[NH_NSCRATCH2] = {};
    double r2dt_tra = 0, r2dt_zdf = 0, r2dt_dyn = 0;   // module r2dt of tra_adv / tra_zdf / dyn_zdf
    int spg_rot_u[3] = {0, 1, 2};                       // dyn_spg_ts level buffers: [now, before, before-before]
    int spg_rot_s[4] = {0, 1, 2, 3};                    // ... and sea level: [now, before, before-before, after]
    bool have_eos = false;                              // eos_init state (nemo_hip_eos_register)
    SeosPar eos{};
    int neos = NEMO_NP_SEOS;
    double *stp_buf = nullptr;                          // stp_ctl partial maxima and results
    double *gsum_buf = nullptr;                         // glob_sum partial pairs and total
    unsigned *spg_bar = nullptr;                        // dyn_spg_ts grid barrier and status
    bool spg_check = false;                             // a cooperative phase 2 awaits its status check
    double *hsb_buf = nullptr;                          // dia_hsb partial pairs and sums
    std::vector<PlanRec> trace;                         // the launches traced for a plan
};
static std::vector<Ctx *> g_ctxs;
static Ctx *g_c = nullptr;   // the current context
// plans: one traced step of every member, replayed as batched launches
struct PlanOp {
    std::string name, kname;
    BatchFn bfn = nullptr;                 // batched: one launch over the members' work list
    unsigned nblk = 0;
    dim3 block;
    int nm = 0;
    size_t off_start = 0, off_ent = 0;     // byte offsets of start[] and the entries in the plan's table
    std::vector<std::function<hipError_t(hipStream_t)>> singles;   // otherwise: one launch per member
};
struct Plan {
    std::vector<PlanOp> ops;
    char *dev = nullptr;                   // the device table of every batched launch
    int nmembers = 0, ngroups = 0, nbatched = 0, nlaunch = 0;
};
static std::vector<Plan *> g_plans;
static void plan_clear_traces()
{
    for (Ctx *c : g_ctxs)
        if (c) c->trace.clear();
}

// recording of a graph on the library stream (nemo_hip_graph_begin / end)
static bool g_capturing = false;
// tracing of a plan (nemo_hip_plan_begin / end): launches are recorded per
// context, nothing is issued
static bool g_planning = false;
static bool recording() { return g_capturing || g_planning; }
static std::vector<hipGraphExec_t> g_graphs;
// staging pool (host-pointer mode)
static std::vector<std::pair<void *, size_t>> g_pool;
// timing: every timed launch records its own start and end events (taken by
// the dispatch packet, so launch gaps are not counted) and adds the elapsed
// time to the routine's total and to the "routine/kernel" total
struct Pending {
    std::string name, key;
    hipEvent_t a, b;
};
static bool g_timing = false;
static std::vector<Pending> g_pend;
static std::vector<hipEvent_t> g_evfree;
static std::map<std::string, std::pair<double, int64_t>> g_tot;
static std::string g_only;   // when set, only launches with this "routine/kernel" key are timed
static std::map<void *, size_t> g_fields;   // resident fields of field_alloc: address -> bytes

static void release_pending()
{
    for (auto &p : g_pend) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        for (const std::string *n : {&p.name, &p.key}) {
            auto &t = g_tot[*n];
            t.first += ms;
            t.second += 1;
        }
        g_evfree.push_back(p.a);
        g_evfree.push_back(p.b);
    }
    g_pend.clear();
}

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
hipStream_t stream() { return g_stream; }
const DevDomain &dom() { return g_c->dom; }
bool capturing() { return recording(); }
bool planning() { return g_planning; }
int plan_add(PlanRec &&r)
{
    g_c->trace.push_back(std::move(r));
    return 0;
}
// tuning switches, process-wide; names and defaults
static const char *g_tune_names[T_NTUNE] = {"fct_fused", "spg_persistent", "zdf_reg", "fct_vlim"};
static int g_tune[T_NTUNE] = {0, 0, 0, 0};
static const int g_tune_max[T_NTUNE] = {2, 2, 1, 2};
int tune(Tune which) { return g_tune[which]; }
int *spg_rot_u() { return g_c->spg_rot_u; }
int *spg_rot_s() { return g_c->spg_rot_s; }
unsigned *spg_barrier() { return g_c->spg_bar; }
void spg_loop_launched() { g_c->spg_check = true; }
void spg_ts_reset()
{
    for (int q = 0; q < 3; q++) g_c->spg_rot_u[q] = q;
    for (int q = 0; q < 4; q++) g_c->spg_rot_s[q] = q;
}
bool have_domain() { return g_c->have_domain; }
void *scratch3(int which) { return g_c->scratch3[which]; }
double *scratch2(int which) { return g_c->scratch2[which]; }

static hipEvent_t ev_get()
{
    if (!g_evfree.empty()) {
        hipEvent_t e = g_evfree.back();
        g_evfree.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
// timers are kept per routine ("tra_adv_fct") and per kernel of the routine
// ("tra_adv_fct/k_fct_low"); the kernel text is the launch expression with
// its template arguments and parentheses removed
static std::string kernel_key(const char *name, const char *kernel)
{
    std::string k;
    int depth = 0;
    for (const char *c = kernel; *c; c++) {
        if (*c == '<') depth++;
        else if (*c == '>') depth--;
        else if (depth == 0 && *c != '(' && *c != ')' && *c != ' ') k += *c;
    }
    return std::string(name) + "/" + k;
}
bool timing_pair(const char *name, const char *kernel, hipEvent_t *start, hipEvent_t *stop)
{
    if (!g_timing || recording()) return false;
    std::string key = kernel_key(name, kernel);
    if (!g_only.empty() && key != g_only) return false;
    if (g_pend.size() >= 4096) {
        // fold early so the event pool stays bounded
        if (hipStreamSynchronize(g_stream) != hipSuccess) return false;
        release_pending();
    }
    hipEvent_t a = ev_get(), b = ev_get();
    if (!a || !b) return false;
    g_pend.push_back({std::string(name), key, a, b});
    *start = a;
    *stop = b;
    return true;
}
static int fold_timing()
{
    HIPCHK(hipStreamSynchronize(g_stream));
    release_pending();
    return 0;
}

static int dev_alloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    HIPCHK(hipMalloc(p, bytes));
    g_c->dom_allocs.push_back(*p);
    return 0;
}
static void free_domain()
{
    for (void *p : g_c->dom_allocs) (void)hipFree(p);
    g_c->dom_allocs.clear();
    for (auto &s : g_c->scratch3) s = nullptr;
    for (auto &s : g_c->scratch2) s = nullptr;
    g_c->stp_buf = nullptr;
    g_c->gsum_buf = g_c->hsb_buf = nullptr;
    g_c->spg_bar = nullptr;
    g_c->spg_check = false;
    g_c->have_domain = false;
}

// ---------------------------------------------------------------------------
// staging of host arrays (NEMO_HIP_PTR_HOST)
// ---------------------------------------------------------------------------
struct Stager {
    struct A {
        void *user, *dev;
        size_t bytes;
        bool out;
    };
    std::vector<A> args;
    int status = 0;
    size_t slot = 0;
    bool host = (g_ptr_mode == NEMO_HIP_PTR_HOST);

    void *get(const void *user, size_t bytes, bool in, bool out)
    {
        if (!user || !host || status) return (void *)user;
        if (recording()) {
            status = set_error(NEMO_HIP_ERR_STATE, "host-pointer calls cannot be recorded into a graph or a plan");
            return nullptr;
        }
        if (slot >= g_pool.size()) g_pool.push_back({nullptr, 0});
        auto &pb = g_pool[slot++];
        if (pb.second < bytes) {
            if (pb.first) (void)hipFree(pb.first);
            pb.first = nullptr;
            pb.second = 0;
            if (hipMalloc(&pb.first, bytes) != hipSuccess) {
                status = set_error(NEMO_HIP_ERR_HIP, "staging allocation of %zu bytes failed", bytes);
                return nullptr;
            }
            pb.second = bytes;
        }
        if (in && hipMemcpyAsync(pb.first, user, bytes, hipMemcpyHostToDevice, g_stream) != hipSuccess)
            status = set_error(NEMO_HIP_ERR_HIP, "host to device copy failed");
        args.push_back({(void *)user, pb.first, bytes, out});
        return pb.first;
    }
    template <typename T> const T *in(const T *p, size_t n) { return (const T *)get(p, n * sizeof(T), true, false); }
    template <typename T> T *inout(T *p, size_t n) { return (T *)get(p, n * sizeof(T), true, true); }
    template <typename T> T *out(T *p, size_t n) { return (T *)get(p, n * sizeof(T), false, true); }

    // copies outputs back (host mode); rc is the status of the launches
    int finish(int rc)
    {
        if (status) return status;
        if (rc) return rc;
        if (!host) return 0;
        for (auto &a : args)
            if (a.out) HIPCHK(hipMemcpyAsync(a.user, a.dev, a.bytes, hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        return 0;
    }
};

// ---------------------------------------------------------------------------
// domain registration
// ---------------------------------------------------------------------------
struct Ival {
    int k0, k1;
};
static Ival inter(Ival a, Ival b)
{
    Ival r{a.k0 > b.k0 ? a.k0 : b.k0, a.k1 < b.k1 ? a.k1 : b.k1};
    if (r.k1 < r.k0) r = Ival{1, 0};
    return r;
}

template <typename V>
static void lbc_host(std::vector<V> &a, int jpi, int jpj, int iper, int jper, char nat, V pval)
{
    auto at = [&](int i, int j) -> V & { return a[(size_t)(j - 1) * jpi + (i - 1)]; };
    for (int j = 1; j <= jpj; j++) {
        if (iper) {
            V x = at(jpi - 1, j), y = at(2, j);
            at(1, j) = x;
            at(jpi, j) = y;
        } else {
            if (nat != 'F') at(1, j) = pval;
            at(jpi, j) = pval;
        }
    }
    for (int i = 1; i <= jpi; i++) {
        if (jper) {
            V x = at(i, jpj - 1), y = at(i, 2);
            at(i, 1) = x;
            at(i, jpj) = y;
        } else {
            if (nat != 'F') at(i, 1) = pval;
            at(i, jpj) = pval;
        }
    }
}

template <typename T>
static int upload(const T *src, size_t n, const T **dst)
{
    void *p;
    int rc = dev_alloc(&p, n * sizeof(T));
    if (rc) return rc;
    HIPCHK(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = (const T *)p;
    return 0;
}
static int upload_f(const double *src, size_t n, const float **dst)
{
    std::vector<float> f(n);
    for (size_t q = 0; q < n; q++) f[q] = (float)src[q];
    return upload(f.data(), n, dst);
}

static int check_options(const nemo_hip_domain_t *h)
{
    if (h->jpi < 4 || h->jpj < 4 || h->jpk < 3) return set_error(NEMO_HIP_ERR_ARG, "domain too small: %d x %d x %d", h->jpi, h->jpj, h->jpk);
    // mask intervals are packed as bytes
    if (h->jpk > 255) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "jpk = %d > 255", h->jpk);
    // kernels address one 3D array (one tracer of a 4D array) with 32-bit byte offsets
    if ((uint64_t)h->jpi * (uint64_t)h->jpj * (uint64_t)h->jpk * sizeof(double) >= (1ull << 32))
        return set_error(NEMO_HIP_ERR_UNSUPPORTED, "domain too large: one 3D FP64 array must stay below 4 GiB");
    if (!(h->jperio == 0 || h->jperio == 1 || h->jperio == 2 || h->jperio == 7))
        return set_error(NEMO_HIP_ERR_UNSUPPORTED, "jperio = %d (north folds are not implemented)", h->jperio);
    if (!h->ln_linssh) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "ln_linssh = F (variable volume) is not implemented");
    if (h->ln_isfcav) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "ln_isfcav = T is not implemented");
    if (h->ln_zps) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "ln_zps = T (partial steps) is not implemented");
    if (h->ln_sco) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "ln_sco = T is not implemented");
    if (h->ln_zad_Aimp) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "ln_zad_Aimp = T is not implemented");
    if (h->nvor_scheme < 0 || h->nvor_scheme > 5) return set_error(NEMO_HIP_ERR_ARG, "nvor_scheme = %d", h->nvor_scheme);
    if (h->n_dynadv < 0 || h->n_dynadv > 2)
        return set_error(NEMO_HIP_ERR_UNSUPPORTED, "n_dynadv = %d (UBS flux form is not implemented)", h->n_dynadv);
    if (h->nn_dynkeg != 0 && h->nn_dynkeg != 1) return set_error(NEMO_HIP_ERR_ARG, "nn_dynkeg = %d", h->nn_dynkeg);
    if (h->nn_een_e3f != 0 && h->nn_een_e3f != 1) return set_error(NEMO_HIP_ERR_ARG, "nn_een_e3f = %d", h->nn_een_e3f);
    if (h->rn_bt_alpha != 0.0) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "rn_bt_alpha /= 0 is not implemented");
    if (h->nn_baro < 1 || h->nn_baro > 300) return set_error(NEMO_HIP_ERR_ARG, "nn_baro = %d", h->nn_baro);
    const void *req[] = {h->k_top, h->k_bot, h->e1t, h->e2t, h->e1u, h->e2u, h->e1v, h->e2v, h->e1e2t, h->r1_e1e2t,
                         h->e1e2u, h->r1_e1e2u, h->e1e2v, h->r1_e1e2v, h->r1_e1e2f, h->e2_e1u, h->e1_e2v, h->r1_e1u,
                         h->r1_e2u, h->r1_e1v, h->r1_e2v, h->ff_f, h->ff_t, h->e3t_n, h->e3u_n, h->e3v_n, h->e3w_n,
                         h->e3f_n, h->e3uw_n, h->e3vw_n};
    for (const void *p : req)
        if (!p) return set_error(NEMO_HIP_ERR_ARG, "domain: a required array pointer is NULL");
    return 0;
}

static int register_domain(const nemo_hip_domain_t *h)
{
    int rc = check_options(h);
    if (rc) return rc;
    free_domain();
    g_c->hdom = *h;
    DevDomain d;
    memset(&d, 0, sizeof d);
    const int jpi = h->jpi, jpj = h->jpj, jpk = h->jpk;
    d.jpi = jpi; d.jpj = jpj; d.jpk = jpk;
    d.iper = (h->jperio == 1 || h->jperio == 7);
    d.jper = (h->jperio == 2 || h->jperio == 7);
    d.n2 = (size_t)jpi * jpj;
    d.n3 = d.n2 * jpk;
    d.ln_traldf_lap = h->ln_traldf_lap; d.ln_traldf_msc = h->ln_traldf_msc; d.ln_zdfddm = h->ln_zdfddm;
    d.n_dynadv = h->n_dynadv; d.nn_dynkeg = h->nn_dynkeg; d.nvor_scheme = h->nvor_scheme; d.nn_een_e3f = h->nn_een_e3f;
    d.ln_drgimp = h->ln_drgimp; d.ln_dynspg_ts = h->ln_dynspg_ts; d.ln_bt_fw = h->ln_bt_fw; d.ln_bt_av = h->ln_bt_av;
    d.nn_bt_flt = h->nn_bt_flt; d.nn_baro = h->nn_baro; d.neuler = h->neuler; d.nit000 = h->nit000;
    d.ln_traqsr = h->ln_traqsr != 0;
    d.rdt = h->rdt; d.rn_atfp = h->rn_atfp; d.rau0 = h->rau0; d.grav = h->grav; d.rn_scal_load = h->rn_scal_load;
    const size_t n2 = d.n2, n3 = d.n3;
    auto I2 = [&](int i, int j) { return (size_t)(j - 1) * jpi + (i - 1); };

    // masks as level intervals: DOM/dommsk.F90:134-197 (tmask, lbc, u/v/f products, lbc)
    std::vector<Ival> T(n2), U(n2, Ival{1, 0}), V(n2, Ival{1, 0}), F(n2, Ival{1, 0});
    for (size_t q = 0; q < n2; q++) {
        const int kt = h->k_top[q], kb = h->k_bot[q];
        if (kt < 0 || kb > jpk || (kt != 0 && kb < kt && kb != 0))
            return set_error(NEMO_HIP_ERR_ARG, "k_top/k_bot out of range at column %zu: %d %d", q, kt, kb);
        Ival r = (kt != 0) ? Ival{kt, kb} : Ival{1, 0};
        if (r.k0 < 1) r.k0 = 1;
        if (r.k1 > jpk) r.k1 = jpk;
        if (r.k1 < r.k0) r = Ival{1, 0};
        T[q] = r;
    }
    lbc_host(T, jpi, jpj, d.iper, d.jper, 'T', Ival{1, 0});
    for (int j = 1; j <= jpj - 1; j++)
        for (int i = 1; i <= jpi - 1; i++) {
            U[I2(i, j)] = inter(T[I2(i, j)], T[I2(i + 1, j)]);
            V[I2(i, j)] = inter(T[I2(i, j)], T[I2(i, j + 1)]);
            F[I2(i, j)] = inter(inter(T[I2(i, j)], T[I2(i + 1, j)]), inter(T[I2(i, j + 1)], T[I2(i + 1, j + 1)]));
        }
    lbc_host(U, jpi, jpj, d.iper, d.jper, 'U', Ival{1, 0});
    lbc_host(V, jpi, jpj, d.iper, d.jper, 'V', Ival{1, 0});
    lbc_host(F, jpi, jpj, d.iper, d.jper, 'F', Ival{1, 0});
    auto up_ival = [&](const std::vector<Ival> &v, KRange &kr) -> int {
        std::vector<int32_t> a(n2), b(n2);
        for (size_t q = 0; q < n2; q++) { a[q] = v[q].k0; b[q] = v[q].k1; }
        int r = upload(a.data(), n2, &kr.k0);
        return r ? r : upload(b.data(), n2, &kr.k1);
    };
    if ((rc = up_ival(T, d.kT)) || (rc = up_ival(U, d.kU)) || (rc = up_ival(V, d.kV)) || (rc = up_ival(F, d.kF))) return rc;
    {
        std::vector<uint32_t> tu(n2), vf(n2);
        auto b = [](int x) { return (uint32_t)(x & 255); };
        for (size_t q = 0; q < n2; q++) {
            tu[q] = b(T[q].k0) | b(T[q].k1) << 8 | b(U[q].k0) << 16 | b(U[q].k1) << 24;
            vf[q] = b(V[q].k0) | b(V[q].k1) << 8 | b(F[q].k0) << 16 | b(F[q].k1) << 24;
        }
        if ((rc = upload(tu.data(), n2, &d.ktu)) || (rc = upload(vf.data(), n2, &d.kvf))) return rc;
    }
    {
        std::vector<double> s(n2), su(n2), sv(n2);
        for (size_t q = 0; q < n2; q++) {
            s[q] = T[q].k1 >= T[q].k0 ? 1.0 : 0.0;
            su[q] = U[q].k1 >= U[q].k0 ? 1.0 : 0.0;
            sv[q] = V[q].k1 >= V[q].k0 ? 1.0 : 0.0;
        }
        if ((rc = upload(s.data(), n2, &d.ssmask)) || (rc = upload(su.data(), n2, &d.ssumask)) ||
            (rc = upload(sv.data(), n2, &d.ssvmask)) || (rc = upload_f(s.data(), n2, &d.ssmask_f)) ||
            (rc = upload_f(su.data(), n2, &d.ssumask_f)) || (rc = upload_f(sv.data(), n2, &d.ssvmask_f)))
            return rc;
    }
    // k-indices, DOM/domzgr.F90:292-314
    {
        std::vector<int32_t> mikt(n2), mbkt(n2);
        std::vector<double> zu(n2, 0.0), zv(n2, 0.0), zbu(n2, 0.0), zbv(n2, 0.0);
        for (size_t q = 0; q < n2; q++) {
            mikt[q] = h->k_top[q] > 1 ? h->k_top[q] : 1;
            mbkt[q] = h->k_bot[q] > 1 ? h->k_bot[q] : 1;
        }
        for (int j = 1; j <= jpj - 1; j++)
            for (int i = 1; i <= jpi - 1; i++) {
                zu[I2(i, j)] = std::max(mikt[I2(i + 1, j)], mikt[I2(i, j)]);
                zv[I2(i, j)] = std::max(mikt[I2(i, j + 1)], mikt[I2(i, j)]);
                zbu[I2(i, j)] = std::min(mbkt[I2(i + 1, j)], mbkt[I2(i, j)]);
                zbv[I2(i, j)] = std::min(mbkt[I2(i, j + 1)], mbkt[I2(i, j)]);
            }
        lbc_host(zu, jpi, jpj, d.iper, d.jper, 'U', 0.0);
        lbc_host(zv, jpi, jpj, d.iper, d.jper, 'V', 0.0);
        lbc_host(zbu, jpi, jpj, d.iper, d.jper, 'U', 0.0);
        lbc_host(zbv, jpi, jpj, d.iper, d.jper, 'V', 0.0);
        std::vector<int32_t> miku(n2), mikv(n2), mbku(n2), mbkv(n2);
        for (size_t q = 0; q < n2; q++) {
            miku[q] = std::max((int)std::lround(zu[q]), 1);
            mikv[q] = std::max((int)std::lround(zv[q]), 1);
            mbku[q] = std::max((int)std::lround(zbu[q]), 1);
            mbkv[q] = std::max((int)std::lround(zbv[q]), 1);
        }
        if ((rc = upload(mikt.data(), n2, &d.mikt)) || (rc = upload(mbkt.data(), n2, &d.mbkt)) ||
            (rc = upload(miku.data(), n2, &d.miku)) || (rc = upload(mikv.data(), n2, &d.mikv)) ||
            (rc = upload(mbku.data(), n2, &d.mbku)) || (rc = upload(mbkv.data(), n2, &d.mbkv)))
            return rc;
    }
    // horizontal metrics and the metric-term pre-computations of DYN/dynvor.F90:919-939
    struct M2 {
        const double *src;
        const double **dd;
        const float **df;
    };
    std::vector<double> di_e2v(n2, 0.0), dj_e1u(n2, 0.0), di_e2u(n2, 0.0), dj_e1v(n2, 0.0);
    for (int j = 2; j <= jpj - 1; j++)
        for (int i = 2; i <= jpi - 1; i++) {
            di_e2u[I2(i, j)] = (h->e2u[I2(i, j)] - h->e2u[I2(i - 1, j)]) * 0.5;
            dj_e1v[I2(i, j)] = (h->e1v[I2(i, j)] - h->e1v[I2(i, j - 1)]) * 0.5;
        }
    lbc_host(di_e2u, jpi, jpj, d.iper, d.jper, 'T', 0.0);
    lbc_host(dj_e1v, jpi, jpj, d.iper, d.jper, 'T', 0.0);
    for (int j = 1; j <= jpj - 1; j++)
        for (int i = 1; i <= jpi - 1; i++) {
            di_e2v[I2(i, j)] = (h->e2v[I2(i + 1, j)] - h->e2v[I2(i, j)]) * 0.5 * h->r1_e1e2f[I2(i, j)];
            dj_e1u[I2(i, j)] = (h->e1u[I2(i, j + 1)] - h->e1u[I2(i, j)]) * 0.5 * h->r1_e1e2f[I2(i, j)];
        }
    lbc_host(di_e2v, jpi, jpj, d.iper, d.jper, 'F', 0.0);
    lbc_host(dj_e1u, jpi, jpj, d.iper, d.jper, 'F', 0.0);
    M2 m2[] = {{h->e1t, &d.md.e1t, &d.mf.e1t}, {h->e2t, &d.md.e2t, &d.mf.e2t}, {h->e1u, &d.md.e1u, &d.mf.e1u},
               {h->e2u, &d.md.e2u, &d.mf.e2u}, {h->e1v, &d.md.e1v, &d.mf.e1v}, {h->e2v, &d.md.e2v, &d.mf.e2v},
               {h->e1e2t, &d.md.e1e2t, &d.mf.e1e2t}, {h->r1_e1e2t, &d.md.r1_e1e2t, &d.mf.r1_e1e2t},
               {h->e1e2u, &d.md.e1e2u, &d.mf.e1e2u}, {h->r1_e1e2u, &d.md.r1_e1e2u, &d.mf.r1_e1e2u},
               {h->e1e2v, &d.md.e1e2v, &d.mf.e1e2v}, {h->r1_e1e2v, &d.md.r1_e1e2v, &d.mf.r1_e1e2v},
               {h->r1_e1e2f, &d.md.r1_e1e2f, &d.mf.r1_e1e2f}, {h->e2_e1u, &d.md.e2_e1u, &d.mf.e2_e1u},
               {h->e1_e2v, &d.md.e1_e2v, &d.mf.e1_e2v}, {h->r1_e1u, &d.md.r1_e1u, &d.mf.r1_e1u},
               {h->r1_e2u, &d.md.r1_e2u, &d.mf.r1_e2u}, {h->r1_e1v, &d.md.r1_e1v, &d.mf.r1_e1v},
               {h->r1_e2v, &d.md.r1_e2v, &d.mf.r1_e2v}, {h->ff_f, &d.md.ff_f, &d.mf.ff_f},
               {h->ff_t, &d.md.ff_t, &d.mf.ff_t}, {di_e2v.data(), &d.md.di_e2v_2e1e2f, &d.mf.di_e2v_2e1e2f},
               {dj_e1u.data(), &d.md.dj_e1u_2e1e2f, &d.mf.dj_e1u_2e1e2f}, {di_e2u.data(), &d.md.di_e2u_2, &d.mf.di_e2u_2},
               {dj_e1v.data(), &d.md.dj_e1v_2, &d.mf.dj_e1v_2}};
    for (auto &e : m2)
        if ((rc = upload(e.src, n2, e.dd)) || (rc = upload_f(e.src, n2, e.df))) return rc;
    // vertical scale factors: a jpk table when horizontally uniform, else 3D
    const double *e3s[7] = {h->e3t_n, h->e3u_n, h->e3v_n, h->e3w_n, h->e3f_n, h->e3uw_n, h->e3vw_n};
    bool uniform = true;
    for (int a = 0; a < 7 && uniform; a++)
        for (int k = 0; k < jpk && uniform; k++) {
            const double v = e3s[a][(size_t)k * n2];
            for (size_t q = 1; q < n2; q++)
                if (e3s[a][(size_t)k * n2 + q] != v) { uniform = false; break; }
        }
    const double **e3d[7] = {&d.md.e3t, &d.md.e3u, &d.md.e3v, &d.md.e3w, &d.md.e3f, &d.md.e3uw, &d.md.e3vw};
    const float **e3f[7] = {&d.mf.e3t, &d.mf.e3u, &d.mf.e3v, &d.mf.e3w, &d.mf.e3f, &d.mf.e3uw, &d.mf.e3vw};
    if (uniform) {
        d.e3_sk = 1;
        d.e3_sij = 0;
        for (int a = 0; a < 7; a++) {
            std::vector<double> col(jpk);
            for (int k = 0; k < jpk; k++) col[k] = e3s[a][(size_t)k * n2];
            if ((rc = upload(col.data(), (size_t)jpk, e3d[a])) || (rc = upload_f(col.data(), (size_t)jpk, e3f[a]))) return rc;
        }
    } else {
        d.e3_sk = n2;
        d.e3_sij = 1;
        for (int a = 0; a < 7; a++)
            if ((rc = upload(e3s[a], n3, e3d[a])) || (rc = upload_f(e3s[a], n3, e3f[a]))) return rc;
    }
    // scratch: one allocation, slots of n3 doubles back to back, so a pair of
    // consecutive slots holds a two-tracer array of either storage type
    {
        void *base;
        if ((rc = dev_alloc(&base, (size_t)S_NSCRATCH3 * n3 * sizeof(double)))) return rc;
        for (int s = 0; s < S_NSCRATCH3; s++) g_c->scratch3[s] = (double *)base + (size_t)s * n3;
        HIPCHK(hipMemset(base, 0, (size_t)S_NSCRATCH3 * n3 * sizeof(double)));
    }
    for (int s = 0; s < NH_NSCRATCH2; s++)
        if ((rc = dev_alloc((void **)&g_c->scratch2[s], n2 * sizeof(double)))) return rc;
    for (int s = 0; s < NH_NSCRATCH2; s++) HIPCHK(hipMemset(g_c->scratch2[s], 0, n2 * sizeof(double)));
    if ((rc = dev_alloc((void **)&g_c->stp_buf, stp_ctl_buffer_bytes()))) return rc;
    if ((rc = dev_alloc((void **)&g_c->gsum_buf, glob_sum_scratch_bytes()))) return rc;
    if ((rc = dev_alloc((void **)&g_c->spg_bar, 64))) return rc;
    HIPCHK(hipMemset(g_c->spg_bar, 0, 64));
    if ((rc = dev_alloc((void **)&g_c->hsb_buf, dia_hsb_scratch_bytes() + 16 * sizeof(double)))) return rc;
    HIPCHK(hipDeviceSynchronize());
    g_c->dom = d;
    g_c->have_domain = true;
    g_c->r2dt_tra = g_c->r2dt_zdf = g_c->r2dt_dyn = 2.0 * h->rdt;
    spg_ts_reset();
    return 0;
}

static double r2dt_update(double &r2dt, int kt)
{
    // tra_adv / tra_zdf / dyn_zdf time step, TRA/traadv.F90:95-97
    if (g_c->hdom.neuler == 0 && kt == g_c->hdom.nit000) r2dt = g_c->hdom.rdt;
    else if (kt <= g_c->hdom.nit000 + 1) r2dt = 2.0 * g_c->hdom.rdt;
    return r2dt;
}
}  // namespace nh

using namespace nh;

#define NEED_DOMAIN()                                                                       \
    do {                                                                                    \
        if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called"); \
        if (!g_c->have_domain) return set_error(NEMO_HIP_ERR_STATE, "no domain registered");     \
    } while (0)
// calls recorded into a graph replay with the same arguments and time step:
// the first steps (Euler step and the switch to leapfrog) are refused
#define NEED_STEADY(kt)                                                                                   \
    do {                                                                                                  \
        if (recording() && (kt) <= g_c->hdom.nit000 + 1)                                                   \
            return set_error(NEMO_HIP_ERR_STATE, "kt = %d: the first two steps cannot be recorded into a graph or a plan", (kt)); \
    } while (0)
#define NEED_NOT_CAPTURING(who)                                                                           \
    do {                                                                                                  \
        if (recording()) return set_error(NEMO_HIP_ERR_STATE, "%s cannot be called while recording a graph or a plan", who); \
    } while (0)
#define NEED_PTR(p)                                                                   \
    do {                                                                              \
        if (!(p)) return set_error(NEMO_HIP_ERR_ARG, "%s: argument %s is NULL", __func__, #p); \
    } while (0)

static int is_tra_type(const char *cdtype, int *is_tra)
{
    if (!cdtype) return set_error(NEMO_HIP_ERR_ARG, "cdtype is NULL");
    if (!strncmp(cdtype, "TRA", 3)) *is_tra = 1;
    else if (!strncmp(cdtype, "TRC", 3)) *is_tra = 0;
    else return set_error(NEMO_HIP_ERR_ARG, "cdtype must be 'TRA' or 'TRC'");
    return 0;
}

extern "C" {

int nemo_hip_init(int device)
{
    if (g_init) {
        if (device == g_device) return NEMO_HIP_OK;
        return set_error(NEMO_HIP_ERR_STATE, "already initialised on device %d", g_device);
    }
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(NEMO_HIP_ERR_ARG, "device %d out of range (%d devices)", device, n);
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreate(&g_stream));
    {
        const int rc = glob_sum_scratch_init(device);   // nemo_hip_glob_sum_noalloc's array
        if (rc) return rc;
    }
    g_device = device;
    g_ctxs.assign(1, new Ctx());
    g_c = g_ctxs[0];
    g_init = true;
    return NEMO_HIP_OK;
}

int nemo_hip_finalize(void)
{
    if (!g_init) return NEMO_HIP_OK;
    if (g_capturing) {
        hipGraph_t gr = nullptr;
        (void)hipStreamEndCapture(g_stream, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        g_capturing = false;
    }
    g_planning = false;
    (void)hipStreamSynchronize(g_stream);
    for (auto &e : g_graphs)
        if (e) (void)hipGraphExecDestroy(e);
    g_graphs.clear();
    for (Plan *P : g_plans) {
        if (!P) continue;
        if (P->dev) (void)hipFree(P->dev);
        delete P;
    }
    g_plans.clear();
    for (Ctx *c : g_ctxs) {
        if (!c) continue;
        g_c = c;
        free_domain();
        delete c;
    }
    g_ctxs.clear();
    g_c = nullptr;
    for (auto &p : g_pool) if (p.first) (void)hipFree(p.first);
    g_pool.clear();
    for (auto &f : g_fields) (void)hipFree(f.first);
    g_fields.clear();
    release_pending();
    for (auto e : g_evfree) (void)hipEventDestroy(e);
    g_evfree.clear();
    (void)hipStreamDestroy(g_stream);
    g_stream = nullptr;
    g_init = false;
    g_device = -1;
    return NEMO_HIP_OK;
}

int nemo_hip_last_error(char *buf, int len)
{
    if (!buf || len <= 0) return NEMO_HIP_ERR_ARG;
    snprintf(buf, (size_t)len, "%s", g_err);
    return NEMO_HIP_OK;
}

int nemo_hip_set_pointer_mode(int mode)
{
    if (mode != NEMO_HIP_PTR_HOST && mode != NEMO_HIP_PTR_DEVICE) return set_error(NEMO_HIP_ERR_ARG, "pointer mode %d", mode);
    g_ptr_mode = mode;
    return NEMO_HIP_OK;
}

int nemo_hip_sync(void)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "not initialised");
    NEED_NOT_CAPTURING("nemo_hip_sync");
    HIPCHK(hipStreamSynchronize(g_stream));
    // a grid barrier of dyn_spg_ts's cooperative phase 2 that gave up leaves
    // its status word set: the barotropic results of that call are wrong
    for (Ctx *c : g_ctxs) {
        if (!c || !c->spg_check || !c->spg_bar) continue;
        c->spg_check = false;
        int flag = 0;
        HIPCHK(hipMemcpy(&flag, c->spg_bar + 2, sizeof flag, hipMemcpyDeviceToHost));
        if (flag) {
            HIPCHK(hipMemset(c->spg_bar, 0, 64));
            return set_error(NEMO_HIP_ERR_HIP, "dyn_spg_ts: a grid barrier of the barotropic sub-steps timed out");
        }
    }
    return NEMO_HIP_OK;
}

void *nemo_hip_stream(void) { return (void *)g_stream; }

int nemo_hip_field_alloc(size_t bytes, void **dev)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "not initialised");
    NEED_NOT_CAPTURING("nemo_hip_field_alloc");
    if (!dev) return set_error(NEMO_HIP_ERR_ARG, "field_alloc: NULL result pointer");
    *dev = nullptr;
    if (bytes == 0) return set_error(NEMO_HIP_ERR_ARG, "field_alloc: zero bytes");
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, bytes));
    g_fields[p] = bytes;
    *dev = p;
    return NEMO_HIP_OK;
}
int nemo_hip_field_free(void *dev)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "not initialised");
    NEED_NOT_CAPTURING("nemo_hip_field_free");
    auto it = g_fields.find(dev);
    if (it == g_fields.end()) return set_error(NEMO_HIP_ERR_ARG, "field_free: %p was not returned by field_alloc", dev);
    HIPCHK(hipStreamSynchronize(g_stream));
    g_fields.erase(it);
    HIPCHK(hipFree(dev));
    return NEMO_HIP_OK;
}
// the field alloc returned that contains [p, p + bytes), or an error
static int field_check(const void *p, size_t bytes, const char *who)
{
    if (!p) return set_error(NEMO_HIP_ERR_ARG, "%s: NULL pointer", who);
    auto it = g_fields.upper_bound(const_cast<void *>(p));
    if (it == g_fields.begin()) return set_error(NEMO_HIP_ERR_ARG, "%s: %p is not a field of field_alloc", who, p);
    --it;
    const char *b = (const char *)it->first, *q = (const char *)p;
    if (q >= b + it->second) return set_error(NEMO_HIP_ERR_ARG, "%s: %p is not a field of field_alloc", who, p);
    if (q + bytes > b + it->second)
        return set_error(NEMO_HIP_ERR_ARG, "%s: %zu bytes at %p exceed the field at %p (%zu bytes)", who, bytes, p,
                         it->first, it->second);
    return NEMO_HIP_OK;
}
int nemo_hip_field_push(void *dev, const void *host, size_t bytes)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "not initialised");
    NEED_NOT_CAPTURING("nemo_hip_field_push");
    if (!host) return set_error(NEMO_HIP_ERR_ARG, "field_push: NULL host pointer");
    int rc = field_check(dev, bytes, "field_push");
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return NEMO_HIP_OK;
}
int nemo_hip_field_pull(void *host, const void *dev, size_t bytes)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "not initialised");
    NEED_NOT_CAPTURING("nemo_hip_field_pull");
    if (!host) return set_error(NEMO_HIP_ERR_ARG, "field_pull: NULL host pointer");
    int rc = field_check(dev, bytes, "field_pull");
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return NEMO_HIP_OK;
}

int nemo_hip_domain_register(const nemo_hip_domain_t *dom)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    NEED_NOT_CAPTURING("nemo_hip_domain_register");
    if (!dom) return set_error(NEMO_HIP_ERR_ARG, "domain is NULL");
    HIPCHK(hipStreamSynchronize(g_stream));
    return register_domain(dom);
}

// ---------------------------------------------------------------------------
// contexts
int nemo_hip_context_create(int *ctx)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    NEED_NOT_CAPTURING("nemo_hip_context_create");
    if (!ctx) return set_error(NEMO_HIP_ERR_ARG, "context_create: NULL result pointer");
    for (size_t q = 1; q < g_ctxs.size(); q++)
        if (!g_ctxs[q]) {
            g_ctxs[q] = new Ctx();
            *ctx = (int)q;
            return NEMO_HIP_OK;
        }
    g_ctxs.push_back(new Ctx());
    *ctx = (int)g_ctxs.size() - 1;
    return NEMO_HIP_OK;
}
int nemo_hip_context_select(int ctx)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    if (ctx < 0 || ctx >= (int)g_ctxs.size() || !g_ctxs[ctx]) return set_error(NEMO_HIP_ERR_ARG, "no context %d", ctx);
    g_c = g_ctxs[ctx];
    return NEMO_HIP_OK;
}
int nemo_hip_context_destroy(int ctx)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    NEED_NOT_CAPTURING("nemo_hip_context_destroy");
    if (ctx <= 0 || ctx >= (int)g_ctxs.size() || !g_ctxs[ctx])
        return set_error(NEMO_HIP_ERR_ARG, "context %d cannot be destroyed (context 0 lives until finalize)", ctx);
    if (g_ctxs[ctx] == g_c) return set_error(NEMO_HIP_ERR_STATE, "context %d is the current context", ctx);
    HIPCHK(hipStreamSynchronize(g_stream));
    Ctx *keep = g_c;
    g_c = g_ctxs[ctx];
    free_domain();
    delete g_c;
    g_ctxs[ctx] = nullptr;
    g_c = keep;
    return NEMO_HIP_OK;
}

// ---------------------------------------------------------------------------
// graphs: the launches of device-mode calls recorded on the library stream
int nemo_hip_graph_begin(void)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    if (recording()) return set_error(NEMO_HIP_ERR_STATE, "a graph or a plan is already being recorded");
    if (g_ptr_mode != NEMO_HIP_PTR_DEVICE) return set_error(NEMO_HIP_ERR_STATE, "graphs record device-pointer calls only");
    HIPCHK(hipStreamSynchronize(g_stream));
    HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    g_capturing = true;
    return NEMO_HIP_OK;
}
int nemo_hip_graph_end(int *graph)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    if (!g_capturing) return set_error(NEMO_HIP_ERR_STATE, "no graph is being recorded");
    g_capturing = false;
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamEndCapture(g_stream, &gr));
    if (!graph) {
        (void)hipGraphDestroy(gr);
        return set_error(NEMO_HIP_ERR_ARG, "graph_end: NULL result pointer");
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return set_error(NEMO_HIP_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
    g_graphs.push_back(ex);
    *graph = (int)g_graphs.size() - 1;
    return NEMO_HIP_OK;
}
int nemo_hip_graph_launch(int graph)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    NEED_NOT_CAPTURING("nemo_hip_graph_launch");
    if (graph < 0 || graph >= (int)g_graphs.size() || !g_graphs[graph]) return set_error(NEMO_HIP_ERR_ARG, "no graph %d", graph);
    HIPCHK(hipGraphLaunch(g_graphs[graph], g_stream));
    return NEMO_HIP_OK;
}
int nemo_hip_graph_free(int graph)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    NEED_NOT_CAPTURING("nemo_hip_graph_free");
    if (graph < 0 || graph >= (int)g_graphs.size() || !g_graphs[graph]) return set_error(NEMO_HIP_ERR_ARG, "no graph %d", graph);
    HIPCHK(hipStreamSynchronize(g_stream));
    HIPCHK(hipGraphExecDestroy(g_graphs[graph]));
    g_graphs[graph] = nullptr;
    return NEMO_HIP_OK;
}

// ---------------------------------------------------------------------------
// plans: one traced step of every member, replayed as batched launches
// ---------------------------------------------------------------------------
int nemo_hip_plan_begin(void)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    if (recording()) return set_error(NEMO_HIP_ERR_STATE, "a graph or a plan is already being recorded");
    if (g_ptr_mode != NEMO_HIP_PTR_DEVICE) return set_error(NEMO_HIP_ERR_STATE, "plans trace device-pointer calls only");
    plan_clear_traces();
    g_planning = true;
    return NEMO_HIP_OK;
}

int nemo_hip_plan_end(int *plan)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    if (!g_planning) return set_error(NEMO_HIP_ERR_STATE, "no plan is being traced");
    g_planning = false;
    if (!plan) {
        plan_clear_traces();
        return set_error(NEMO_HIP_ERR_ARG, "plan_end: NULL result pointer");
    }
    // members: the contexts that traced launches; members whose launch
    // sequences match (same kernels, argument types and workgroup shapes)
    // form a group, and each position of a group's sequence becomes one op
    std::vector<Ctx *> mem;
    for (Ctx *c : g_ctxs)
        if (c && !c->trace.empty()) mem.push_back(c);
    if (mem.empty()) return set_error(NEMO_HIP_ERR_STATE, "plan_end: no launch was traced");
    std::vector<std::pair<std::string, std::vector<Ctx *>>> groups;
    for (Ctx *c : mem) {
        std::string key;
        char buf[96];
        for (const PlanRec &r : c->trace) {
            snprintf(buf, sizeof buf, "|%p:%u,%u,%u:%zu:", (const void *)r.bfn, r.block.x, r.block.y, r.block.z,
                     r.ent.size());
            key += r.kname;
            key += buf;
        }
        size_t gq = 0;
        while (gq < groups.size() && groups[gq].first != key) gq++;
        if (gq == groups.size()) groups.push_back({key, {}});
        groups[gq].second.push_back(c);
    }
    Plan *P = new Plan();
    P->nmembers = (int)mem.size();
    P->ngroups = (int)groups.size();
    std::vector<char> tab;
    auto align = [&](size_t a) {
        while (tab.size() % a) tab.push_back(0);
    };
    for (auto &gr : groups) {
        const std::vector<Ctx *> &gm = gr.second;
        const int nm = (int)gm.size();
        const size_t n = gm[0]->trace.size();
        for (size_t q = 0; q < n; q++) {
            const PlanRec &r0 = gm[0]->trace[q];
            PlanOp op;
            op.name = r0.routine;
            op.kname = r0.kname;
            if (r0.bfn) {
                std::vector<unsigned> st((size_t)nm + 1);
                unsigned acc = 0;
                for (int m = 0; m < nm; m++) {
                    const PlanRec &r = gm[m]->trace[q];
                    st[m] = acc;
                    const unsigned long long nb = (unsigned long long)r.grid.x * r.grid.y * r.grid.z;
                    const unsigned long long padded = (nb + 7ull) & ~7ull;
                    if (acc + padded > 0x7fffffffull) {
                        delete P;
                        plan_clear_traces();
                        return set_error(NEMO_HIP_ERR_UNSUPPORTED, "plan_end: %s: more than 2^31 workgroups",
                                         r0.kname.c_str());
                    }
                    acc += (unsigned)padded;
                }
                st[nm] = acc;
                align(16);
                op.off_start = tab.size();
                tab.insert(tab.end(), (const char *)st.data(), (const char *)(st.data() + st.size()));
                align(16);
                op.off_ent = tab.size();
                for (int m = 0; m < nm; m++) {
                    const std::vector<char> &e = gm[m]->trace[q].ent;
                    tab.insert(tab.end(), e.begin(), e.end());
                }
                op.bfn = r0.bfn;
                op.nblk = acc;
                op.block = r0.block;
                op.nm = nm;
                P->nbatched++;
                P->nlaunch++;
            } else {
                for (int m = 0; m < nm; m++) op.singles.push_back(gm[m]->trace[q].single);
                P->nlaunch += nm;
            }
            P->ops.push_back(std::move(op));
        }
    }
    plan_clear_traces();
    if (!tab.empty()) {
        if (hipMalloc(&P->dev, tab.size()) != hipSuccess ||
            hipMemcpy(P->dev, tab.data(), tab.size(), hipMemcpyHostToDevice) != hipSuccess) {
            if (P->dev) (void)hipFree(P->dev);
            delete P;
            return set_error(NEMO_HIP_ERR_HIP, "plan_end: the plan's table of %zu bytes could not be placed",
                             tab.size());
        }
    }
    g_plans.push_back(P);
    *plan = (int)g_plans.size() - 1;
    return NEMO_HIP_OK;
}

int nemo_hip_plan_launch(int plan)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    if (g_planning) return set_error(NEMO_HIP_ERR_STATE, "plan_launch while a plan is traced");
    if (plan < 0 || plan >= (int)g_plans.size() || !g_plans[plan]) return set_error(NEMO_HIP_ERR_ARG, "no plan %d", plan);
    const Plan *P = g_plans[plan];
    for (const PlanOp &op : P->ops) {
        if (op.bfn) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (!timing_pair(op.name.c_str(), op.kname.c_str(), &e0, &e1)) e0 = e1 = nullptr;
            HIPCHK(op.bfn(op.nblk, op.block, g_stream, e0, e1, (const unsigned *)(P->dev + op.off_start), op.nm,
                          P->dev + op.off_ent));
        } else {
            for (const auto &f : op.singles) HIPCHK(f(g_stream));
        }
    }
    return NEMO_HIP_OK;
}

int nemo_hip_plan_info(int plan, int *nmembers, int *ngroups, int *nops, int *nbatched, int *nlaunch)
{
    if (plan < 0 || plan >= (int)g_plans.size() || !g_plans[plan]) return set_error(NEMO_HIP_ERR_ARG, "no plan %d", plan);
    const Plan *P = g_plans[plan];
    if (nmembers) *nmembers = P->nmembers;
    if (ngroups) *ngroups = P->ngroups;
    if (nops) *nops = (int)P->ops.size();
    if (nbatched) *nbatched = P->nbatched;
    if (nlaunch) *nlaunch = P->nlaunch;
    return NEMO_HIP_OK;
}

int nemo_hip_plan_free(int plan)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    NEED_NOT_CAPTURING("nemo_hip_plan_free");
    if (plan < 0 || plan >= (int)g_plans.size() || !g_plans[plan]) return set_error(NEMO_HIP_ERR_ARG, "no plan %d", plan);
    HIPCHK(hipStreamSynchronize(g_stream));
    Plan *P = g_plans[plan];
    if (P->dev) HIPCHK(hipFree(P->dev));
    delete P;
    g_plans[plan] = nullptr;
    return NEMO_HIP_OK;
}

int nemo_hip_timing_enable(int on)
{
    g_timing = on != 0;
    return NEMO_HIP_OK;
}
int nemo_hip_timing_only(const char *key)
{
    g_only = key ? key : "";
    return NEMO_HIP_OK;
}
int nemo_hip_timing_reset(void)
{
    if (g_init) {
        int rc = fold_timing();
        if (rc) return rc;
    }
    g_tot.clear();
    return NEMO_HIP_OK;
}
int nemo_hip_timing_get(const char *name, double *total_ms, int64_t *count)
{
    if (!name || !total_ms || !count) return set_error(NEMO_HIP_ERR_ARG, "timing_get: NULL argument");
    if (g_init) {
        int rc = fold_timing();
        if (rc) return rc;
    }
    auto it = g_tot.find(name);
    *total_ms = it == g_tot.end() ? 0.0 : it->second.first;
    *count = it == g_tot.end() ? 0 : it->second.second;
    return NEMO_HIP_OK;
}

// ---------------------------------------------------------------------------
int nemo_hip_dom_msk(double *tmask, double *umask, double *vmask, double *fmask, double *wmask, double *wumask,
                     double *wvmask, double *ssmask, double *ssumask, double *ssvmask, int32_t *mikt, int32_t *mbkt,
                     int32_t *miku, int32_t *mikv, int32_t *mbku, int32_t *mbkv)
{
    NEED_DOMAIN();
    const DevDomain &d = g_c->dom;
    Stager S;
    double *m3[7] = {S.out(tmask, d.n3), S.out(umask, d.n3), S.out(vmask, d.n3), S.out(fmask, d.n3),
                     S.out(wmask, d.n3), S.out(wumask, d.n3), S.out(wvmask, d.n3)};
    double *m2[3] = {S.out(ssmask, d.n2), S.out(ssumask, d.n2), S.out(ssvmask, d.n2)};
    int32_t *ki[6] = {S.out(mikt, d.n2), S.out(mbkt, d.n2), S.out(miku, d.n2),
                      S.out(mikv, d.n2), S.out(mbku, d.n2), S.out(mbkv, d.n2)};
    const int32_t *src[6] = {d.mikt, d.mbkt, d.miku, d.mikv, d.mbku, d.mbkv};
    int rc = 0;
    if (!S.status) rc = launch_dom_msk(m3[0], m3[1], m3[2], m3[3], m3[4], m3[5], m3[6], m2[0], m2[1], m2[2]);
    for (int q = 0; q < 6 && !rc && !S.status; q++)
        if (ki[q]) HIPCHK(hipMemcpyAsync(ki[q], src[q], d.n2 * sizeof(int32_t), hipMemcpyDeviceToDevice, g_stream));
    return S.finish(rc);
}

int nemo_hip_lbc_lnk(double *ptab, int nk, char cd_nat, double psgn, double pval)
{
    (void)psgn;
    NEED_DOMAIN();
    NEED_PTR(ptab);
    if (nk < 1) return set_error(NEMO_HIP_ERR_ARG, "lbc_lnk: nk = %d", nk);
    if (!strchr("TUVWF", cd_nat)) return set_error(NEMO_HIP_ERR_ARG, "lbc_lnk: grid type '%c'", cd_nat);
    Stager S;
    double *p = S.inout(ptab, g_c->dom.n2 * (size_t)nk);
    int rc = S.status ? S.status : launch_lbc<double>(p, nk, cd_nat, pval);
    return S.finish(rc);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// tracers (templated bodies shared by the FP64 and FP32 entry points)
// ---------------------------------------------------------------------------
template <typename T>
static int do_tra_adv_trp(const T *un, const T *vn, const T *wn, T *zun, T *zvn, T *zwn)
{
    NEED_DOMAIN();
    NEED_PTR(un); NEED_PTR(vn); NEED_PTR(wn); NEED_PTR(zun); NEED_PTR(zvn); NEED_PTR(zwn);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *a = S.in(un, n3), *b = S.in(vn, n3), *c = S.in(wn, n3);
    T *x = S.out(zun, n3), *y = S.out(zvn, n3), *z = S.out(zwn, n3);
    int rc = S.status ? S.status : launch_tra_adv_trp<T>(a, b, c, x, y, z);
    return S.finish(rc);
}

template <typename T>
static int do_tra_adv_fct(int kt, int kit000, const char *cdtype, double p2dt, const T *pun, const T *pvn,
                          const T *pwn, const T *ptb, const T *ptn, T *pta, int kjpt, int kn_fct_h, int kn_fct_v)
{
    (void)kt; (void)kit000;
    NEED_DOMAIN();
    NEED_STEADY(kt);
    int is_tra, rc = is_tra_type(cdtype, &is_tra);
    if (rc) return rc;
    NEED_PTR(pun); NEED_PTR(pvn); NEED_PTR(pwn); NEED_PTR(ptb); NEED_PTR(ptn); NEED_PTR(pta);
    if (kjpt < 1) return set_error(NEMO_HIP_ERR_ARG, "kjpt = %d", kjpt);
    if ((kn_fct_h != 2 && kn_fct_h != 4) || (kn_fct_v != 2 && kn_fct_v != 4))
        return set_error(NEMO_HIP_ERR_UNSUPPORTED, "tra_adv_fct: kn_fct_h = %d, kn_fct_v = %d (2 or 4)", kn_fct_h, kn_fct_v);
    if (!(p2dt > 0.0)) return set_error(NEMO_HIP_ERR_ARG, "p2dt = %g", p2dt);
    const size_t n3 = g_c->dom.n3, n4 = n3 * (size_t)kjpt;
    Stager S;
    const T *u = S.in(pun, n3), *v = S.in(pvn, n3), *w = S.in(pwn, n3), *tb = S.in(ptb, n4), *tn = S.in(ptn, n4);
    T *ta = S.inout(pta, n4);
    rc = S.status ? S.status : launch_tra_adv_fct<T>(p2dt, u, v, w, tb, tn, ta, kjpt, kn_fct_h, kn_fct_v);
    return S.finish(rc);
}

// tra_adv: r2dt of kt, transports formed where used, FCT
template <typename T>
static int do_tra_adv(int kt, const T *un, const T *vn, const T *wn, const T *ptb, const T *ptn, T *pta, int kjpt,
                      int kn_fct_h, int kn_fct_v)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    NEED_PTR(un); NEED_PTR(vn); NEED_PTR(wn); NEED_PTR(ptb); NEED_PTR(ptn); NEED_PTR(pta);
    if (kjpt < 1) return set_error(NEMO_HIP_ERR_ARG, "kjpt = %d", kjpt);
    if ((kn_fct_h != 2 && kn_fct_h != 4) || (kn_fct_v != 2 && kn_fct_v != 4))
        return set_error(NEMO_HIP_ERR_UNSUPPORTED, "tra_adv: kn_fct_h = %d, kn_fct_v = %d (2 or 4)", kn_fct_h, kn_fct_v);
    const double r2dt = r2dt_update(g_c->r2dt_tra, kt);
    const size_t n3 = g_c->dom.n3, n4 = n3 * (size_t)kjpt;
    Stager S;
    const T *u = S.in(un, n3), *v = S.in(vn, n3), *w = S.in(wn, n3), *tb = S.in(ptb, n4), *tn = S.in(ptn, n4);
    T *ta = S.inout(pta, n4);
    int rc = S.status ? S.status : launch_tra_adv<T>(r2dt, u, v, w, tb, tn, ta, kjpt, kn_fct_h, kn_fct_v);
    return S.finish(rc);
}

template <typename T>
static int do_tra_ldf_lap(int kt, int kit000, const char *cdtype, const T *pahu, const T *pahv, const T *ptb, T *pta,
                          int kjpt, int kpass)
{
    (void)kt; (void)kit000;
    NEED_DOMAIN();
    int is_tra, rc = is_tra_type(cdtype, &is_tra);
    if (rc) return rc;
    NEED_PTR(pahu); NEED_PTR(pahv); NEED_PTR(ptb); NEED_PTR(pta);
    if (kjpt < 1 || (kpass != 1 && kpass != 2)) return set_error(NEMO_HIP_ERR_ARG, "kjpt = %d, kpass = %d", kjpt, kpass);
    const size_t n3 = g_c->dom.n3, n4 = n3 * (size_t)kjpt;
    Stager S;
    const T *a = S.in(pahu, n3), *b = S.in(pahv, n3), *tb = S.in(ptb, n4);
    T *ta = S.inout(pta, n4);
    rc = S.status ? S.status : launch_tra_ldf_lap<T>(a, b, tb, ta, kjpt, kpass);
    return S.finish(rc);
}

template <typename T>
static int do_tra_ldf_iso(int kt, int kit000, const char *cdtype, const T *pahu, const T *pahv, const T *ptb,
                          const T *ptbb, T *pta, int kjpt, int kpass, const T *uslp, const T *vslp, const T *wslpi,
                          const T *wslpj, T *ah_wslp2, T *akz)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    if (recording() && kt == kit000)
        return set_error(NEMO_HIP_ERR_STATE, "tra_ldf_iso: the first call (kt = kit000) cannot be recorded into a graph or a plan");
    int is_tra, rc = is_tra_type(cdtype, &is_tra);
    if (rc) return rc;
    NEED_PTR(pahu); NEED_PTR(pahv); NEED_PTR(ptb); NEED_PTR(pta); NEED_PTR(uslp); NEED_PTR(vslp);
    NEED_PTR(wslpi); NEED_PTR(wslpj); NEED_PTR(ah_wslp2); NEED_PTR(akz);
    if (kjpt < 1 || (kpass != 1 && kpass != 2)) return set_error(NEMO_HIP_ERR_ARG, "kjpt = %d, kpass = %d", kjpt, kpass);
    if (kpass == 2 && !ptbb) return set_error(NEMO_HIP_ERR_ARG, "tra_ldf_iso: kpass = 2 needs ptbb");
    if (!ptbb) ptbb = ptb;
    // time step of traldf_iso.F90:131-133
    const double z2dt = (g_c->hdom.neuler == 0 && kt == g_c->hdom.nit000) ? g_c->hdom.rdt : 2.0 * g_c->hdom.rdt;
    const size_t n3 = g_c->dom.n3, n4 = n3 * (size_t)kjpt;
    Stager S;
    const T *a = S.in(pahu, n3), *b = S.in(pahv, n3), *tb = S.in(ptb, n4);
    const T *tbb = (ptbb == ptb) ? tb : S.in(ptbb, n4);
    T *ta = S.inout(pta, n4);
    const T *us = S.in(uslp, n3), *vs = S.in(vslp, n3), *wi = S.in(wslpi, n3), *wj = S.in(wslpj, n3);
    T *ah = S.inout(ah_wslp2, n3), *kz = S.inout(akz, n3);
    rc = S.status ? S.status
                  : launch_tra_ldf_iso<T>(kpass == 1 && kt == kit000, z2dt, a, b, tb, tbb, ta, kjpt, kpass, us, vs,
                                          wi, wj, ah, kz);
    return S.finish(rc);
}

template <typename T>
static int do_tra_zdf_imp(int kt, int kit000, const char *cdtype, double p2dt, const T *ptb, T *pta, int kjpt,
                          const T *avt, const T *avs, int l_ldfslp, const T *ah_wslp2, const T *akz, int clamp_sal)
{
    (void)kt; (void)kit000;
    NEED_DOMAIN();
    NEED_STEADY(kt);
    int is_tra, rc = is_tra_type(cdtype, &is_tra);
    if (rc) return rc;
    NEED_PTR(ptb); NEED_PTR(pta);
    if (kjpt < 1) return set_error(NEMO_HIP_ERR_ARG, "kjpt = %d", kjpt);
    if (is_tra && !avt) return set_error(NEMO_HIP_ERR_ARG, "tra_zdf_imp: avt is NULL");
    if ((!is_tra || g_c->dom.ln_zdfddm) && !avs) return set_error(NEMO_HIP_ERR_ARG, "tra_zdf_imp: avs is NULL");
    const T *add_user = g_c->dom.ln_traldf_msc ? akz : ah_wslp2;
    if (l_ldfslp && !add_user) return set_error(NEMO_HIP_ERR_ARG, "tra_zdf_imp: iso-neutral coefficient is NULL");
    const size_t n3 = g_c->dom.n3, n4 = n3 * (size_t)kjpt;
    Stager S;
    const T *tb = S.in(ptb, n4);
    T *ta = S.inout(pta, n4);
    const T *t = S.in(avt, n3);
    const T *s = (avs == avt) ? t : S.in(avs, n3);
    const T *add = l_ldfslp ? S.in(add_user, n3) : nullptr;
    rc = S.status ? S.status : launch_tra_zdf_imp<T>(p2dt, is_tra, tb, ta, kjpt, t, s, l_ldfslp, add, clamp_sal);
    return S.finish(rc);
}

// ---------------------------------------------------------------------------
// momentum
// ---------------------------------------------------------------------------
template <typename T>
static int do_dyn_adv(int which, int kscheme, const T *un, const T *vn, const T *wn, T *ua, T *va)
{
    NEED_DOMAIN();
    NEED_PTR(un); NEED_PTR(vn); NEED_PTR(ua); NEED_PTR(va);
    if ((which & 2 || which == 4) && !wn) return set_error(NEMO_HIP_ERR_ARG, "wn is NULL");
    if (kscheme != 0 && kscheme != 1) return set_error(NEMO_HIP_ERR_ARG, "kscheme = %d", kscheme);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *u = S.in(un, n3), *v = S.in(vn, n3), *w = wn ? S.in(wn, n3) : nullptr;
    T *a = S.inout(ua, n3), *b = S.inout(va, n3);
    int rc = S.status ? S.status : launch_dyn_adv<T>(which, kscheme, u, v, w, a, b);
    return S.finish(rc);
}

template <typename T>
static int dyn_adv_dispatch(const T *un, const T *vn, const T *wn, T *ua, T *va)
{
    NEED_DOMAIN();
    switch (g_c->dom.n_dynadv) {
    case NEMO_NP_LIN_DYN: return NEMO_HIP_OK;
    case NEMO_NP_VEC_C2: return do_dyn_adv<T>(3, g_c->dom.nn_dynkeg, un, vn, wn, ua, va);
    case NEMO_NP_FLX_C2: return do_dyn_adv<T>(4, 0, un, vn, wn, ua, va);
    default: return set_error(NEMO_HIP_ERR_UNSUPPORTED, "dyn_adv: n_dynadv = %d", g_c->dom.n_dynadv);
    }
}

template <typename T>
static int do_dyn_vor(const T *un, const T *vn, T *ua, T *va)
{
    NEED_DOMAIN();
    NEED_PTR(un); NEED_PTR(vn); NEED_PTR(ua); NEED_PTR(va);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *u = S.in(un, n3), *v = S.in(vn, n3);
    T *a = S.inout(ua, n3), *b = S.inout(va, n3);
    int rc = S.status ? S.status : launch_dyn_vor<T>(u, v, a, b);
    return S.finish(rc);
}

template <typename T>
static int do_dyn_zdf(int kt, const T *ub, const T *vb, T *ua, T *va, const T *avm, const double *rCdU_bot,
                      const double *ua_b, const double *va_b, const double *utau, const double *utau_b,
                      const double *vtau, const double *vtau_b)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    NEED_PTR(ub); NEED_PTR(vb); NEED_PTR(ua); NEED_PTR(va); NEED_PTR(avm); NEED_PTR(rCdU_bot);
    NEED_PTR(utau); NEED_PTR(utau_b); NEED_PTR(vtau); NEED_PTR(vtau_b);
    if (!g_c->dom.ln_drgimp) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "dyn_zdf: explicit drag (ln_drgimp = F) is not implemented");
    if (g_c->dom.ln_dynspg_ts && (!ua_b || !va_b)) return set_error(NEMO_HIP_ERR_ARG, "dyn_zdf: ua_b/va_b are NULL");
    const double r2dt = r2dt_update(g_c->r2dt_dyn, kt);
    const size_t n3 = g_c->dom.n3, n2 = g_c->dom.n2;
    Stager S;
    const T *b1 = S.in(ub, n3), *b2 = S.in(vb, n3), *m = S.in(avm, n3);
    T *a1 = S.inout(ua, n3), *a2 = S.inout(va, n3);
    const double *cd = S.in(rCdU_bot, n2), *uab = S.in(ua_b, n2), *vab = S.in(va_b, n2), *t1 = S.in(utau, n2),
                 *t2 = S.in(utau_b, n2), *t3 = S.in(vtau, n2), *t4 = S.in(vtau_b, n2);
    int rc = S.status ? S.status : launch_dyn_zdf<T>(r2dt, b1, b2, a1, a2, m, cd, uab, vab, t1, t2, t3, t4);
    return S.finish(rc);
}

template <typename T>
static int do_dyn_spg_ts(int kt, nemo_hip_spg_ts_t *st)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    if (!st) return set_error(NEMO_HIP_ERR_ARG, "dyn_spg_ts: state is NULL");
    if (!g_c->dom.ln_dynspg_ts) return set_error(NEMO_HIP_ERR_STATE, "dyn_spg_ts called with ln_dynspg_ts = F");
    const void *req[] = {st->ua, st->va, st->un, st->vn, st->un_b, st->vn_b, st->ub_b, st->vb_b, st->sshn, st->hu_n,
                         st->hv_n, st->r1_hu_n, st->r1_hv_n, st->ht_n, st->rCdU_bot, st->utau, st->vtau, st->emp,
                         st->ua_b, st->va_b, st->ssha, st->un_adv, st->vn_adv, st->ub2_b, st->vb2_b, st->un_bf,
                         st->vn_bf};
    for (const void *p : req)
        if (!p) return set_error(NEMO_HIP_ERR_ARG, "dyn_spg_ts: a required field is NULL");
    const size_t n3 = g_c->dom.n3, n2 = g_c->dom.n2;
    Stager S;
    nemo_hip_spg_ts_t D;
    D.ua = S.inout((T *)st->ua, n3); D.va = S.inout((T *)st->va, n3);
    D.un = S.inout((T *)st->un, n3); D.vn = S.inout((T *)st->vn, n3);
    D.un_b = S.in(st->un_b, n2); D.vn_b = S.in(st->vn_b, n2); D.ub_b = S.in(st->ub_b, n2); D.vb_b = S.in(st->vb_b, n2);
    D.sshn = S.in(st->sshn, n2); D.hu_n = S.in(st->hu_n, n2); D.hv_n = S.in(st->hv_n, n2);
    D.r1_hu_n = S.in(st->r1_hu_n, n2); D.r1_hv_n = S.in(st->r1_hv_n, n2); D.ht_n = S.in(st->ht_n, n2);
    D.rCdU_bot = S.in(st->rCdU_bot, n2); D.utau = S.in(st->utau, n2); D.vtau = S.in(st->vtau, n2);
    D.emp = S.in(st->emp, n2); D.rnf = st->rnf ? S.in(st->rnf, n2) : nullptr;
    D.ua_b = S.out(st->ua_b, n2); D.va_b = S.out(st->va_b, n2); D.ssha = S.out(st->ssha, n2);
    D.un_adv = S.out(st->un_adv, n2); D.vn_adv = S.out(st->vn_adv, n2);
    D.ub2_b = S.inout(st->ub2_b, n2); D.vb2_b = S.inout(st->vb2_b, n2);
    D.un_bf = S.inout(st->un_bf, n2); D.vn_bf = S.inout(st->vn_bf, n2);
    int rc = S.status ? S.status : launch_dyn_spg_ts<T>(kt, &D);
    return S.finish(rc);
}

// ---------------------------------------------------------------------------
// fused steps (resident fields)
// ---------------------------------------------------------------------------
template <typename T>
static int do_step_dyn(int kt, const T *ub, const T *vb, const T *wn, const T *avm, const double *utau_b,
                       const double *vtau_b, nemo_hip_spg_ts_t *st)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    if (g_ptr_mode != NEMO_HIP_PTR_DEVICE)
        return set_error(NEMO_HIP_ERR_STATE, "step_dyn: the fused step works on device-resident fields only");
    if (!st) return set_error(NEMO_HIP_ERR_ARG, "step_dyn: state is NULL");
    if (!g_c->dom.ln_dynspg_ts) return set_error(NEMO_HIP_ERR_STATE, "step_dyn needs ln_dynspg_ts = T");
    if (!g_c->dom.ln_drgimp) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "step_dyn: explicit drag (ln_drgimp = F) is not implemented");
    if (!g_c->dom.ln_bt_fw) return set_error(NEMO_HIP_ERR_UNSUPPORTED, "step_dyn: ln_bt_fw = F is not implemented");
    NEED_PTR(ub); NEED_PTR(vb); NEED_PTR(avm); NEED_PTR(utau_b); NEED_PTR(vtau_b);
    if (!wn && g_c->dom.n_dynadv != NEMO_NP_LIN_DYN) return set_error(NEMO_HIP_ERR_ARG, "step_dyn: wn is NULL");
    const void *req[] = {st->ua, st->va, st->un, st->vn, st->un_b, st->vn_b, st->ub_b, st->vb_b, st->sshn, st->hu_n,
                         st->hv_n, st->r1_hu_n, st->r1_hv_n, st->ht_n, st->rCdU_bot, st->utau, st->vtau, st->emp,
                         st->ua_b, st->va_b, st->ssha, st->un_adv, st->vn_adv, st->ub2_b, st->vb2_b, st->un_bf,
                         st->vn_bf};
    for (const void *p : req)
        if (!p) return set_error(NEMO_HIP_ERR_ARG, "step_dyn: a required field is NULL");
    const double r2dt = r2dt_update(g_c->r2dt_dyn, kt);
    return launch_step_dyn<T>(kt, r2dt, ub, vb, wn, avm, utau_b, vtau_b, st);
}

template <typename T>
static int do_step_tra(int kt, const T *un, const T *vn, const T *wn, const T *tsb, const T *tsn, T *tsa,
                       int kn_fct_h, int kn_fct_v, const T *ahtu, const T *ahtv, const T *uslp, const T *vslp,
                       const T *wslpi, const T *wslpj, T *ah_wslp2, T *akz, const T *avt)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    if (g_ptr_mode != NEMO_HIP_PTR_DEVICE)
        return set_error(NEMO_HIP_ERR_STATE, "step_tra: the fused step works on device-resident fields only");
    NEED_PTR(un); NEED_PTR(vn); NEED_PTR(wn); NEED_PTR(tsb); NEED_PTR(tsn); NEED_PTR(tsa); NEED_PTR(ahtu);
    NEED_PTR(ahtv); NEED_PTR(uslp); NEED_PTR(vslp); NEED_PTR(wslpi); NEED_PTR(wslpj); NEED_PTR(ah_wslp2);
    NEED_PTR(akz); NEED_PTR(avt);
    if ((kn_fct_h != 2 && kn_fct_h != 4) || (kn_fct_v != 2 && kn_fct_v != 4))
        return set_error(NEMO_HIP_ERR_ARG, "step_tra: kn_fct_h = %d, kn_fct_v = %d", kn_fct_h, kn_fct_v);
    // r2dt of tra_adv and tra_zdf (TRA/traadv.F90:95-97, TRA/trazdf.F90:66-68), z2dt of traldf_iso.F90:131-133
    const double p2dt = r2dt_update(g_c->r2dt_tra, kt);
    const double p2dt_zdf = r2dt_update(g_c->r2dt_zdf, kt);
    const double z2dt = (g_c->hdom.neuler == 0 && kt == g_c->hdom.nit000) ? g_c->hdom.rdt : 2.0 * g_c->hdom.rdt;
    return launch_step_tra<T>(p2dt, z2dt, p2dt_zdf, un, vn, wn, tsb, tsn, tsa, kn_fct_h, kn_fct_v, ahtu, ahtv, uslp,
                              vslp, wslpi, wslpj, avt, ah_wslp2, akz);
}

// ---------------------------------------------------------------------------
// time-step closure: div_hor, ssh_nxt, wzv, ssh_swp, tra_nxt, dyn_nxt
// ---------------------------------------------------------------------------
static bool euler_step(int kt) { return g_c->hdom.neuler == 0 && kt == g_c->hdom.nit000; }

template <typename T>
static int do_div_hor(int kt, const T *un, const T *vn, T *hdivn)
{
    (void)kt;
    NEED_DOMAIN();
    NEED_PTR(un); NEED_PTR(vn); NEED_PTR(hdivn);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *u = S.in(un, n3), *v = S.in(vn, n3);
    T *h = S.inout(hdivn, n3);
    int rc = S.status ? S.status : launch_div_hor<T>(u, v, h);
    return S.finish(rc);
}

template <typename T>
static int do_ssh_nxt(int kt, const T *un, const T *vn, T *hdivn, const double *sshb, const double *emp_b,
                      const double *emp, double *ssha)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    NEED_PTR(un); NEED_PTR(vn); NEED_PTR(hdivn); NEED_PTR(sshb); NEED_PTR(emp_b); NEED_PTR(emp); NEED_PTR(ssha);
    // DYN/sshwzv.F90:85-87
    const double z2dt = euler_step(kt) ? g_c->hdom.rdt : 2.0 * g_c->hdom.rdt;
    const double zcoef = 0.5 * (1.0 / g_c->hdom.rau0);
    const size_t n3 = g_c->dom.n3, n2 = g_c->dom.n2;
    Stager S;
    const T *u = S.in(un, n3), *v = S.in(vn, n3);
    T *h = S.inout(hdivn, n3);
    const double *b = S.in(sshb, n2), *e0 = S.in(emp_b, n2), *e1 = S.in(emp, n2);
    double *a = S.out(ssha, n2);
    int rc = S.status ? S.status : launch_ssh_nxt<T>(z2dt, zcoef, u, v, h, b, e0, e1, a);
    return S.finish(rc);
}

template <typename T>
static int do_wzv(int kt, const T *hdivn, T *wn)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    NEED_PTR(hdivn); NEED_PTR(wn);
    // DYN/sshwzv.F90:163-164
    const double z1_2dt = euler_step(kt) ? 1. / g_c->hdom.rdt : 1. / (2. * g_c->hdom.rdt);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *h = S.in(hdivn, n3);
    T *w = S.inout(wn, n3);
    int rc = S.status ? S.status : launch_wzv<T>(kt == g_c->hdom.nit000, z1_2dt, h, w);
    return S.finish(rc);
}

template <typename T>
static int do_tra_nxt(int kt, T *tsb, T *tsn, T *tsa)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    NEED_PTR(tsb); NEED_PTR(tsn); NEED_PTR(tsa);
    const size_t n4 = g_c->dom.n3 * 2;
    Stager S;
    T *b = S.inout(tsb, n4), *n = S.inout(tsn, n4), *a = S.inout(tsa, n4);
    int rc = S.status ? S.status : launch_tra_nxt<T>(euler_step(kt), b, n, a);
    return S.finish(rc);
}

template <typename T>
static int do_dyn_nxt(int kt, T *ub, T *vb, T *un, T *vn, T *ua, T *va, const double *ua_b, const double *va_b,
                      const double *un_adv, const double *vn_adv, double *ub_b, double *vb_b, double *un_b,
                      double *vn_b, const double *r1_hu_n, const double *r1_hv_n)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    NEED_PTR(ub); NEED_PTR(vb); NEED_PTR(un); NEED_PTR(vn); NEED_PTR(ua); NEED_PTR(va);
    NEED_PTR(ub_b); NEED_PTR(vb_b); NEED_PTR(un_b); NEED_PTR(vn_b); NEED_PTR(r1_hu_n); NEED_PTR(r1_hv_n);
    const int ts = g_c->dom.ln_dynspg_ts;
    if (ts && (!ua_b || !va_b)) return set_error(NEMO_HIP_ERR_ARG, "dyn_nxt: ua_b/va_b are NULL");
    if (ts && !g_c->dom.ln_bt_fw && (!un_adv || !vn_adv))
        return set_error(NEMO_HIP_ERR_ARG, "dyn_nxt: un_adv/vn_adv are NULL (ln_bt_fw = F)");
    const size_t n3 = g_c->dom.n3, n2 = g_c->dom.n2;
    Stager S;
    T *b1 = S.inout(ub, n3), *b2 = S.inout(vb, n3), *n1 = S.inout(un, n3), *n2_ = S.inout(vn, n3);
    T *a1 = S.inout(ua, n3), *a2 = S.inout(va, n3);
    const double *uab = ts ? S.in(ua_b, n2) : nullptr, *vab = ts ? S.in(va_b, n2) : nullptr;
    const bool adv = ts && !g_c->dom.ln_bt_fw;
    const double *uad = adv ? S.in(un_adv, n2) : nullptr, *vad = adv ? S.in(vn_adv, n2) : nullptr;
    double *ubb = S.inout(ub_b, n2), *vbb = S.inout(vb_b, n2), *unb = S.inout(un_b, n2), *vnb = S.inout(vn_b, n2);
    const double *rh = S.in(r1_hu_n, n2), *rv = S.in(r1_hv_n, n2);
    int rc = S.status ? S.status
                      : launch_dyn_nxt<T>(euler_step(kt), b1, b2, n1, n2_, a1, a2, uab, vab, uad, vad, ubb, vbb, unb,
                                          vnb, rh, rv);
    return S.finish(rc);
}

// ---------------------------------------------------------------------------
// density, pressure gradient, lateral viscosity
// ---------------------------------------------------------------------------
#define NEED_EOS()                                                                                         \
    do {                                                                                                   \
        if (!g_c->have_eos) return set_error(NEMO_HIP_ERR_STATE, "no equation of state registered (eos_register)"); \
    } while (0)

template <typename T>
static int do_eos(int pot, const T *pts, T *prd, T *prhop, const T *pdep)
{
    NEED_DOMAIN();
    NEED_EOS();
    NEED_PTR(pts); NEED_PTR(prd); NEED_PTR(pdep);
    if (pot && !prhop) return set_error(NEMO_HIP_ERR_ARG, "eos_insitu_pot: prhop is NULL");
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *ts = S.in(pts, 2 * n3), *dep = S.in(pdep, n3);
    T *rd = S.inout(prd, n3), *rp = pot ? S.inout(prhop, n3) : nullptr;
    int rc = S.status ? S.status : launch_eos<T>(g_c->neos, g_c->eos, pot, ts, rd, rp, dep);
    return S.finish(rc);
}

template <typename T>
static int do_rab(const T *pts, T *pab, const T *gdept)
{
    NEED_DOMAIN();
    NEED_EOS();
    NEED_PTR(pts); NEED_PTR(pab); NEED_PTR(gdept);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *ts = S.in(pts, 2 * n3), *gt = S.in(gdept, n3);
    T *ab = S.inout(pab, 2 * n3);
    int rc = S.status ? S.status : launch_rab<T>(g_c->neos, g_c->eos, ts, ab, gt);
    return S.finish(rc);
}

template <typename T>
static int do_bn2(const T *pts, const T *pab, T *pn2, const T *gdept, const T *gdepw)
{
    NEED_DOMAIN();
    NEED_PTR(pts); NEED_PTR(pab); NEED_PTR(pn2); NEED_PTR(gdept); NEED_PTR(gdepw);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *ts = S.in(pts, 2 * n3), *ab = S.in(pab, 2 * n3), *gt = S.in(gdept, n3), *gw = S.in(gdepw, n3);
    T *n2 = S.inout(pn2, n3);
    int rc = S.status ? S.status : launch_bn2<T>(ts, ab, n2, gt, gw);
    return S.finish(rc);
}

template <typename T>
static int do_hpg_zco(int kt, const T *rhd, T *ua, T *va)
{
    (void)kt;
    NEED_DOMAIN();
    NEED_PTR(rhd); NEED_PTR(ua); NEED_PTR(va);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *r = S.in(rhd, n3);
    T *a = S.inout(ua, n3), *b = S.inout(va, n3);
    int rc = S.status ? S.status : launch_hpg_zco<T>(r, a, b);
    return S.finish(rc);
}

template <typename T>
static int do_dyn_ldf_lap(int kt, const T *pub, const T *pvb, T *pua, T *pva, int kpass, const T *ahmt,
                          const T *ahmf)
{
    (void)kt;
    NEED_DOMAIN();
    NEED_PTR(pub); NEED_PTR(pvb); NEED_PTR(pua); NEED_PTR(pva); NEED_PTR(ahmt); NEED_PTR(ahmf);
    if (kpass != 1 && kpass != 2) return set_error(NEMO_HIP_ERR_ARG, "dyn_ldf_lap: kpass = %d", kpass);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const T *ub = S.in(pub, n3), *vb = S.in(pvb, n3), *at = S.in(ahmt, n3), *af = S.in(ahmf, n3);
    T *a = S.inout(pua, n3), *b = S.inout(pva, n3);
    int rc = S.status ? S.status : launch_dyn_ldf_lap<T>(kpass == 1 ? 1.0 : -1.0, ub, vb, a, b, at, af);
    return S.finish(rc);
}

// zdf_mxl (nmln, hmlp, hmlpt) and ldf_slp
template <typename T>
static int do_zdf_mxl(int kt, const T *rn2b, const T *gdept, const T *gdepw, int nlb10, double rho_c,
                      int32_t *nmln, double *hmlp, double *hmlpt)
{
    (void)kt;
    NEED_DOMAIN();
    NEED_PTR(rn2b); NEED_PTR(gdept); NEED_PTR(gdepw); NEED_PTR(nmln); NEED_PTR(hmlp); NEED_PTR(hmlpt);
    if (nlb10 < 2 || nlb10 > g_c->dom.jpk - 1)
        return set_error(NEMO_HIP_ERR_ARG, "zdf_mxl: nlb10 = %d outside 2..jpk-1 = %d", nlb10, g_c->dom.jpk - 1);
    const size_t n2 = g_c->dom.n2, n3 = g_c->dom.n3;
    Stager S;
    const T *r = S.in(rn2b, n3), *gt = S.in(gdept, n3), *gw = S.in(gdepw, n3);
    int32_t *nm = S.out(nmln, n2);
    double *h1 = S.out(hmlp, n2), *h2 = S.out(hmlpt, n2);
    int rc = S.status ? S.status : launch_zdf_mxl<T>(r, gt, gw, nlb10, rho_c, nm, h1, h2);
    return S.finish(rc);
}

template <typename T>
static int do_ldf_slp(int kt, const T *prd, const T *pn2, const T *gdept, const T *gdepw, const int32_t *nmln,
                      const double *hmlp, const double *hmlpt, double rn_slpmax, T *uslp, T *vslp, T *wslpi,
                      T *wslpj)
{
    (void)kt;
    NEED_DOMAIN();
    NEED_PTR(prd); NEED_PTR(pn2); NEED_PTR(gdept); NEED_PTR(gdepw); NEED_PTR(nmln); NEED_PTR(hmlp);
    NEED_PTR(hmlpt); NEED_PTR(uslp); NEED_PTR(vslp); NEED_PTR(wslpi); NEED_PTR(wslpj);
    if (!(rn_slpmax > 0.)) return set_error(NEMO_HIP_ERR_ARG, "ldf_slp: rn_slpmax = %g", rn_slpmax);
    const size_t n2 = g_c->dom.n2, n3 = g_c->dom.n3;
    Stager S;
    const T *rd = S.in(prd, n3), *n2p = S.in(pn2, n3), *gt = S.in(gdept, n3), *gw = S.in(gdepw, n3);
    const int32_t *nm = S.in(nmln, n2);
    const double *h1 = S.in(hmlp, n2), *h2 = S.in(hmlpt, n2);
    T *us = S.inout(uslp, n3), *vs = S.inout(vslp, n3), *wi = S.inout(wslpi, n3), *wj = S.inout(wslpj, n3);
    int rc = S.status ? S.status : launch_ldf_slp<T>(rd, n2p, gt, gw, nm, h1, h2, rn_slpmax, us, vs, wi, wj);
    return S.finish(rc);
}

// tra_sbc (linear free surface, no penetrative solar flux)
template <typename T>
static int do_tra_sbc(int kt, double *qns, double *qsr, const double *sfx, const double *emp, const T *tsn, T *tsa,
                      double *sbc, double *sbc_b)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    NEED_PTR(qns); NEED_PTR(qsr); NEED_PTR(sfx); NEED_PTR(emp); NEED_PTR(tsn); NEED_PTR(tsa); NEED_PTR(sbc);
    NEED_PTR(sbc_b);
    const size_t n2 = g_c->dom.n2, n3 = g_c->dom.n3;
    Stager S;
    double *q1 = S.inout(qns, n2), *q2 = S.inout(qsr, n2), *c = S.inout(sbc, 2 * n2), *cb = S.inout(sbc_b, 2 * n2);
    const double *sf = S.in(sfx, n2), *em = S.in(emp, n2);
    const T *tn = S.in(tsn, 2 * n3);
    T *ta = S.inout(tsa, 2 * n3);
    int rc = S.status ? S.status : launch_tra_sbc<T>(kt == g_c->hdom.nit000, q1, q2, sf, em, tn, ta, c, cb);
    return S.finish(rc);
}

// tra_qsr (two bands)
template <typename T>
static int do_tra_qsr(int kt, int kqsr_b, int nksr, double rn_abs, double rn_si0, double rn_si1, const double *qsr,
                      const T *gdepw, double *qsr_hc, double *qsr_hc_b, T *tsa, double *fraqsr_1lev)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    NEED_PTR(qsr); NEED_PTR(gdepw); NEED_PTR(qsr_hc); NEED_PTR(qsr_hc_b); NEED_PTR(tsa); NEED_PTR(fraqsr_1lev);
    if (nksr < 1 || nksr > g_c->dom.jpk - 1)
        return set_error(NEMO_HIP_ERR_ARG, "tra_qsr: nksr = %d outside 1..jpk-1", nksr);
    if (!(rn_si0 > 0.0) || !(rn_si1 > 0.0))
        return set_error(NEMO_HIP_ERR_ARG, "tra_qsr: extinction depths must be positive");
    const size_t n2 = g_c->dom.n2, n3 = g_c->dom.n3;
    const int mode = kt != g_c->hdom.nit000 ? 0 : (kqsr_b ? 2 : 1);
    Stager S;
    const double *q = S.in(qsr, n2);
    const T *gw = S.in(gdepw, n3);
    double *hc = S.inout(qsr_hc, n3), *hb = S.inout(qsr_hc_b, n3), *fr = S.inout(fraqsr_1lev, n2);
    T *ta = S.inout(tsa, 2 * n3);
    int rc = S.status ? S.status : launch_tra_qsr<T>(mode, nksr, rn_abs, rn_si0, rn_si1, q, gw, hc, hb, ta, fr);
    return S.finish(rc);
}

// glob_sum / glob_sum_full: the compensated sum into psum (one double)
template <typename T>
static int do_glob_sum(const T *ptab, int nk, int kind, double *psum)
{
    NEED_DOMAIN();
    NEED_PTR(ptab); NEED_PTR(psum);
    const DevDomain &d = g_c->dom;
    if (nk < 1 || nk > d.jpk) return set_error(NEMO_HIP_ERR_ARG, "glob_sum: nk = %d outside 1..jpk", nk);
    if (kind != 0 && kind != 1) return set_error(NEMO_HIP_ERR_ARG, "glob_sum: kind = %d (0 tmask_i, 1 tmask_h)", kind);
    Stager S;
    const T *p = S.in(ptab, d.n2 * (size_t)nk);
    double *o = S.out(psum, 1);
    int rc = S.status ? S.status : launch_glob_sum<T>(p, nk, kind, g_c->gsum_buf, o);
    return S.finish(rc);
}

// dia_hsb: the thirteen sums on the device, the budget on the host
static const double HSB_RCP = 3991.86795711963;   // TRA/eosbn2.F90:1245
template <typename T>
static int do_dia_hsb_init(nemo_hip_hsb_t *h)
{
    NEED_DOMAIN();
    NEED_NOT_CAPTURING("dia_hsb_init");
    NEED_PTR(h);
    NEED_PTR(h->tsn); NEED_PTR(h->sshn); NEED_PTR(h->surf_ini); NEED_PTR(h->ssh_ini); NEED_PTR(h->ssh_hc_loc_ini);
    NEED_PTR(h->ssh_sc_loc_ini); NEED_PTR(h->e3t_ini); NEED_PTR(h->hc_loc_ini); NEED_PTR(h->sc_loc_ini);
    const size_t n2 = g_c->dom.n2, n3 = g_c->dom.n3;
    Stager S;
    const double *sn = S.in(h->sshn, n2);
    const T *ts = S.in((const T *)h->tsn, 2 * n3);
    double *si = S.out(h->surf_ini, n2), *shi = S.out(h->ssh_ini, n2), *shc = S.out(h->ssh_hc_loc_ini, n2);
    double *ssc = S.out(h->ssh_sc_loc_ini, n2), *e3i = S.out(h->e3t_ini, n3), *hci = S.out(h->hc_loc_ini, n3);
    double *sci = S.out(h->sc_loc_ini, n3);
    double *tot = g_c->hsb_buf + dia_hsb_scratch_bytes() / sizeof(double);
    int rc = S.status ? S.status : launch_dia_hsb_init<T>(sn, ts, si, shi, shc, ssc, e3i, hci, sci);
    // surf_tot = glob_sum(surf), diahsb.F90:412
    if (!rc) rc = launch_glob_sum<double>(si, 1, 0, g_c->gsum_buf, tot);
    if ((rc = S.finish(rc))) return rc;
    HIPCHK(hipMemcpyAsync(&h->surf_tot, tot, sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    h->frc_v = h->frc_t = h->frc_s = h->frc_wn_t = h->frc_wn_s = 0.0;   // diahsb.F90:291-305
    return NEMO_HIP_OK;
}

template <typename T>
static int do_dia_hsb(int kt, nemo_hip_hsb_t *h)
{
    NEED_DOMAIN();
    NEED_NOT_CAPTURING("dia_hsb");
    NEED_PTR(h);
    NEED_PTR(h->tsn); NEED_PTR(h->tsb); NEED_PTR(h->wn); NEED_PTR(h->sshn); NEED_PTR(h->emp); NEED_PTR(h->rnf);
    NEED_PTR(h->sbc_tsc); NEED_PTR(h->surf_ini); NEED_PTR(h->ssh_ini); NEED_PTR(h->ssh_hc_loc_ini);
    NEED_PTR(h->ssh_sc_loc_ini); NEED_PTR(h->e3t_ini); NEED_PTR(h->hc_loc_ini); NEED_PTR(h->sc_loc_ini);
    const int traqsr = g_c->dom.ln_traqsr;
    if (traqsr) NEED_PTR(h->qsr);
    const size_t n2 = g_c->dom.n2, n3 = g_c->dom.n3;
    Stager S;
    const double *em = S.in(h->emp, n2), *rn = S.in(h->rnf, n2), *sb = S.in(h->sbc_tsc, 2 * n2);
    const double *qs = traqsr ? S.in(h->qsr, n2) : em;
    const double *sn = S.in(h->sshn, n2);
    const T *tn = S.in((const T *)h->tsn, 2 * n3), *tb = S.in((const T *)h->tsb, 2 * n3);
    const T *w = S.in((const T *)h->wn, n3);
    const double *si = S.in(h->surf_ini, n2), *shi = S.in(h->ssh_ini, n2), *shc = S.in(h->ssh_hc_loc_ini, n2);
    const double *ssc = S.in(h->ssh_sc_loc_ini, n2), *e3i = S.in(h->e3t_ini, n3), *hci = S.in(h->hc_loc_ini, n3);
    const double *sci = S.in(h->sc_loc_ini, n3);
    double *sums = g_c->hsb_buf + dia_hsb_scratch_bytes() / sizeof(double);
    int rc = S.status ? S.status
                      : launch_dia_hsb<T>(traqsr, em, rn, sb, qs, sn, tn, tb, w, si, shi, shc, ssc, e3i, hci, sci,
                                          g_c->hsb_buf, sums);
    if ((rc = S.finish(rc))) return rc;
    double z[13];
    HIPCHK(hipMemcpyAsync(z, sums, sizeof z, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    for (int q = 0; q < 13; q++) h->sums[q] = z[q];
    // diahsb.F90:93-233, linear free surface
    const double rau0 = g_c->hdom.rau0, rdt = g_c->hdom.rdt;
    const double r1_rau0 = 1.0 / rau0, r1_rau0_rcp = 1.0 / (rau0 * HSB_RCP);
    const double z_frc_trd_v = r1_rau0 * z[0];
    double z_frc_trd_t = z[1];
    const double z_frc_trd_s = z[2];
    if (traqsr) z_frc_trd_t = z_frc_trd_t + r1_rau0_rcp * z[3];
    const double z_wn_trd_t = -z[4], z_wn_trd_s = -z[5];
    h->frc_v = h->frc_v + z_frc_trd_v * rdt;
    h->frc_t = h->frc_t + z_frc_trd_t * rdt;
    h->frc_s = h->frc_s + z_frc_trd_s * rdt;
    h->frc_wn_t = h->frc_wn_t + z_wn_trd_t * rdt;
    h->frc_wn_s = h->frc_wn_s + z_wn_trd_s * rdt;
    const double zdiff_v1 = z[6] - h->frc_v;
    const double z_ssh_hc = z[7], z_ssh_sc = z[8];
    const double zdiff_hc = z[10] - h->frc_t, zdiff_sc = z[11] - h->frc_s;
    const double zdiff_hc1 = zdiff_hc + z_ssh_hc, zdiff_sc1 = zdiff_sc + z_ssh_sc;
    const double zerr_hc1 = z_ssh_hc - h->frc_wn_t, zerr_sc1 = z_ssh_sc - h->frc_wn_s;
    const double zvol_tot = z[12];
    const double ztime = h->surf_tot * (double)kt * rdt;
    h->bgfrcvol = h->frc_v * 1.e-9;
    h->bgfrctem = h->frc_t * rau0 * HSB_RCP * 1.e-20;
    h->bgfrchfx = h->frc_t * rau0 * HSB_RCP / ztime;
    h->bgfrcsal = h->frc_s * 1.e-9;
    h->bgtemper = zdiff_hc1 / zvol_tot;
    h->bgsaline = zdiff_sc1 / zvol_tot;
    h->bgheatco = zdiff_hc1 * 1.e-20 * rau0 * HSB_RCP;
    h->bgheatfx = zdiff_hc1 * rau0 * HSB_RCP / ztime;
    h->bgsaltco = zdiff_sc1 * 1.e-9;
    h->bgvolssh = zdiff_v1 * 1.e-9;
    h->bgmistem = zerr_hc1 / zvol_tot;
    h->bgmissal = zerr_sc1 / zvol_tot;
    return NEMO_HIP_OK;
}

// stp_ctl: maxima on the device, the tests and the message on the host
template <typename T>
static int do_stp_ctl(int kt, const double *sshn, const T *un, const T *tsn, double zmax[6], int32_t loc[11],
                      int *kindic)
{
    NEED_DOMAIN();
    NEED_NOT_CAPTURING("stp_ctl");
    NEED_PTR(sshn); NEED_PTR(un); NEED_PTR(tsn); NEED_PTR(zmax); NEED_PTR(loc); NEED_PTR(kindic);
    const DevDomain &d = g_c->dom;
    Stager S;
    const double *ss = S.in(sshn, d.n2);
    const T *u = S.in(un, d.n3), *ts = S.in(tsn, 2 * d.n3);
    if (S.status) return S.status;
    int rc = launch_stp_ctl<T>(ss, u, ts, g_c->stp_buf, 0);
    if (rc) return rc;
    double *dz = g_c->stp_buf + (size_t)1024 * 6;
    HIPCHK(hipMemcpyAsync(zmax, dz, 6 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    // stpctl.F90:149-155
    const bool bad = zmax[0] > 20. || zmax[1] > 10. || zmax[2] >= 0. || zmax[3] >= 100. || zmax[3] < 0. ||
                     std::isnan(zmax[0] + zmax[1] + zmax[2]);
    if (!bad) return NEMO_HIP_OK;
    if ((rc = launch_stp_ctl<T>(ss, u, ts, g_c->stp_buf, 1))) return rc;
    uint32_t q[4];
    HIPCHK(hipMemcpyAsync(q, (const void *)(dz + 8), sizeof q, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    const size_t n2 = d.n2;
    auto ijk = [&](uint32_t v, int32_t *o, int nd) {
        if (v == 0xffffffffu) { for (int c = 0; c < nd; c++) o[c] = 0; return; }
        o[0] = (int32_t)(v % n2 % (size_t)d.jpi) + 1;
        o[1] = (int32_t)(v % n2 / (size_t)d.jpi) + 1;
        if (nd == 3) o[2] = (int32_t)(v / n2) + 1;
    };
    ijk(q[0], loc, 2);
    ijk(q[1], loc + 2, 3);
    ijk(q[2], loc + 5, 3);
    ijk(q[3], loc + 8, 3);
    *kindic = -3;
    return set_error(NEMO_HIP_ERR_NUMERIC,
                     "stp_ctl: |ssh| > 20 m  or  |U| > 10 m/s  or  S <= 0  or  S >= 100  or  NaN encounter in the tests; "
                     "kt=%d |ssh| max %g at (%d,%d); |U| max %g at (%d,%d,%d); S min %g at (%d,%d,%d); "
                     "S max %g at (%d,%d,%d)", kt, zmax[0], loc[0], loc[1], zmax[1], loc[2], loc[3], loc[4], -zmax[2],
                     loc[5], loc[6], loc[7], zmax[3], loc[8], loc[9], loc[10]);
}

// ---------------------------------------------------------------------------
// FP64 entry points
// ---------------------------------------------------------------------------
extern "C" {
int nemo_hip_tra_adv_trp(const double *un, const double *vn, const double *wn, double *zun, double *zvn, double *zwn)
{
    return do_tra_adv_trp<double>(un, vn, wn, zun, zvn, zwn);
}

int nemo_hip_tra_adv_fct(int kt, int kit000, const char *cdtype, double p2dt, const double *pun, const double *pvn,
                         const double *pwn, const double *ptb, const double *ptn, double *pta, int kjpt, int kn_fct_h,
                         int kn_fct_v)
{
    return do_tra_adv_fct<double>(kt, kit000, cdtype, p2dt, pun, pvn, pwn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v);
}

int nemo_hip_interp_4th_cpt(const double *pt_in, double *pt_out)
{
    NEED_DOMAIN();
    NEED_PTR(pt_in); NEED_PTR(pt_out);
    Stager S;
    const double *a = S.in(pt_in, g_c->dom.n3);
    double *b = S.inout(pt_out, g_c->dom.n3);
    int rc = S.status ? S.status : launch_interp_4th_cpt(a, b);
    return S.finish(rc);
}

int nemo_hip_tra_ldf_lap(int kt, int kit000, const char *cdtype, const double *pahu, const double *pahv,
                         const double *pgu, const double *pgv, const double *pgui, const double *pgvi,
                         const double *ptb, double *pta, int kjpt, int kpass)
{
    (void)pgu; (void)pgv; (void)pgui; (void)pgvi;   // partial-step gradients: ln_zps = F
    return do_tra_ldf_lap<double>(kt, kit000, cdtype, pahu, pahv, ptb, pta, kjpt, kpass);
}

int nemo_hip_tra_ldf_iso(int kt, int kit000, const char *cdtype, const double *pahu, const double *pahv,
                         const double *pgu, const double *pgv, const double *pgui, const double *pgvi,
                         const double *ptb, const double *ptbb, double *pta, int kjpt, int kpass,
                         const double *uslp, const double *vslp, const double *wslpi, const double *wslpj,
                         double *ah_wslp2, double *akz)
{
    (void)pgu; (void)pgv; (void)pgui; (void)pgvi;
    return do_tra_ldf_iso<double>(kt, kit000, cdtype, pahu, pahv, ptb, ptbb, pta, kjpt, kpass, uslp, vslp, wslpi,
                                  wslpj, ah_wslp2, akz);
}

int nemo_hip_tra_zdf_imp(int kt, int kit000, const char *cdtype, double p2dt, const double *ptb, double *pta, int kjpt,
                         const double *avt, const double *avs, int l_ldfslp, const double *ah_wslp2, const double *akz)
{
    return do_tra_zdf_imp<double>(kt, kit000, cdtype, p2dt, ptb, pta, kjpt, avt, avs, l_ldfslp, ah_wslp2, akz, 0);
}

int nemo_hip_tra_zdf(int kt, const double *tsb, double *tsa, const double *avt, const double *avs, int l_ldfslp,
                     const double *ah_wslp2, const double *akz)
{
    NEED_DOMAIN();
    const double r2dt = r2dt_update(g_c->r2dt_zdf, kt);
    return do_tra_zdf_imp<double>(kt, g_c->hdom.nit000, "TRA", r2dt, tsb, tsa, 2, avt, avs, l_ldfslp, ah_wslp2, akz, 1);
}

int nemo_hip_dyn_keg(int kt, int kscheme, const double *un, const double *vn, double *ua, double *va)
{
    (void)kt;
    return do_dyn_adv<double>(1, kscheme, un, vn, nullptr, ua, va);
}
int nemo_hip_dyn_zad(int kt, const double *un, const double *vn, const double *wn, double *ua, double *va)
{
    (void)kt;
    return do_dyn_adv<double>(2, 0, un, vn, wn, ua, va);
}
int nemo_hip_dyn_adv_cen2(int kt, const double *un, const double *vn, const double *wn, double *ua, double *va)
{
    (void)kt;
    return do_dyn_adv<double>(4, 0, un, vn, wn, ua, va);
}
int nemo_hip_dyn_adv(int kt, const double *un, const double *vn, const double *wn, double *ua, double *va)
{
    (void)kt;
    return dyn_adv_dispatch<double>(un, vn, wn, ua, va);
}
int nemo_hip_dyn_vor(int kt, const double *un, const double *vn, double *ua, double *va)
{
    (void)kt;
    return do_dyn_vor<double>(un, vn, ua, va);
}
int nemo_hip_dyn_zdf(int kt, const double *ub, const double *vb, double *ua, double *va, const double *avm,
                     const double *rCdU_bot, const double *ua_b, const double *va_b, const double *utau,
                     const double *utau_b, const double *vtau, const double *vtau_b)
{
    return do_dyn_zdf<double>(kt, ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau, utau_b, vtau, vtau_b);
}
int nemo_hip_dyn_spg_ts(int kt, nemo_hip_spg_ts_t *st) { return do_dyn_spg_ts<double>(kt, st); }

int nemo_hip_step_dyn(int kt, const double *ub, const double *vb, const double *wn, const double *avm,
                      const double *utau_b, const double *vtau_b, nemo_hip_spg_ts_t *st)
{
    return do_step_dyn<double>(kt, ub, vb, wn, avm, utau_b, vtau_b, st);
}

int nemo_hip_step_tra(int kt, const double *un, const double *vn, const double *wn, const double *tsb,
                      const double *tsn, double *tsa, int kn_fct_h, int kn_fct_v, const double *ahtu,
                      const double *ahtv, const double *uslp, const double *vslp, const double *wslpi,
                      const double *wslpj, double *ah_wslp2, double *akz, const double *avt)
{
    return do_step_tra<double>(kt, un, vn, wn, tsb, tsn, tsa, kn_fct_h, kn_fct_v, ahtu, ahtv, uslp, vslp, wslpi,
                               wslpj, ah_wslp2, akz, avt);
}

int nemo_hip_div_hor(int kt, const double *un, const double *vn, double *hdivn)
{
    return do_div_hor<double>(kt, un, vn, hdivn);
}
int nemo_hip_ssh_nxt(int kt, const double *un, const double *vn, double *hdivn, const double *sshb,
                     const double *emp_b, const double *emp, double *ssha)
{
    return do_ssh_nxt<double>(kt, un, vn, hdivn, sshb, emp_b, emp, ssha);
}
int nemo_hip_wzv(int kt, const double *hdivn, double *wn) { return do_wzv<double>(kt, hdivn, wn); }
int nemo_hip_ssh_swp(int kt, double *sshb, double *sshn, const double *ssha)
{
    NEED_DOMAIN();
    NEED_STEADY(kt);
    NEED_PTR(sshb); NEED_PTR(sshn); NEED_PTR(ssha);
    const size_t n2 = g_c->dom.n2;
    Stager S;
    double *b = S.inout(sshb, n2), *n = S.inout(sshn, n2);
    const double *a = S.in(ssha, n2);
    int rc = S.status ? S.status : launch_ssh_swp(euler_step(kt), b, n, a);
    return S.finish(rc);
}
int nemo_hip_tra_nxt(int kt, double *tsb, double *tsn, double *tsa) { return do_tra_nxt<double>(kt, tsb, tsn, tsa); }
int nemo_hip_dyn_nxt(int kt, double *ub, double *vb, double *un, double *vn, double *ua, double *va,
                     const double *ua_b, const double *va_b, const double *un_adv, const double *vn_adv,
                     double *ub_b, double *vb_b, double *un_b, double *vn_b, const double *r1_hu_n,
                     const double *r1_hv_n)
{
    return do_dyn_nxt<double>(kt, ub, vb, un, vn, ua, va, ua_b, va_b, un_adv, vn_adv, ub_b, vb_b, un_b, vn_b,
                              r1_hu_n, r1_hv_n);
}

// ---------------------------------------------------------------------------
// restart files (rst_write, rst_read, ts_rst): a directory holding
// manifest.json and one raw Fortran-ordered array per variable.  The
// variable names are the reference's restart names.
// ---------------------------------------------------------------------------
namespace {

struct RstVar {
    const char *name;
    void *ptr;      // caller's pointer, host or device per the pointer mode
    size_t n;       // elements
    int width;      // bytes per element in the caller's array
    bool before;    // a before field (absent from a file: Euler restart)
    bool ts;        // a ts_rst field
    bool opt;       // optional: written when given, read when given and present
};

// reading: the destinations of the variables; writing: their sources (the
// solar heat content is written from qsr_hc under the name qsr_hc_b, the
// reference's convention, traqsr.F90:291)
std::vector<RstVar> rst_vars(const nemo_hip_rst_t *r, bool reading = false)
{
    const size_t n2 = g_c->dom.n2, n3 = g_c->dom.n3;
    const int w = r->real_bytes;
    auto tr = [&](void *p, int t) { return p ? (void *)((char *)p + (size_t)t * n3 * (size_t)w) : nullptr; };
    return {{"ub", r->ub, n3, w, true, false},          {"vb", r->vb, n3, w, true, false},
            {"tb", tr(r->tsb, 0), n3, w, true, false},  {"sb", tr(r->tsb, 1), n3, w, true, false},
            {"sshb", r->sshb, n2, 8, true, false},      {"un", r->un, n3, w, false, false},
            {"vn", r->vn, n3, w, false, false},         {"tn", tr(r->tsn, 0), n3, w, false, false},
            {"sn", tr(r->tsn, 1), n3, w, false, false}, {"sshn", r->sshn, n2, 8, false, false},
            {"rhop", r->rhop, n3, w, false, false},     {"ub2_b", r->ub2_b, n2, 8, false, true},
            {"vb2_b", r->vb2_b, n2, 8, false, true},    {"un_bf", r->un_bf, n2, 8, false, true},
            {"vn_bf", r->vn_bf, n2, 8, false, true},
            {"qsr_hc_b", reading ? (void *)r->qsr_hc_b : (void *)r->qsr_hc, n3, 8, false, false, true},
            {"fraqsr_1lev", r->fraqsr_1lev, n2, 8, false, false, true}};
}

int rst_check(const nemo_hip_rst_t *r, const char *who, bool reading)
{
    if (r->real_bytes != 8 && r->real_bytes != 4)
        return set_error(NEMO_HIP_ERR_ARG, "%s: real_bytes = %d (8 or 4)", who, r->real_bytes);
    if (!g_c->dom.ln_bt_av)
        return set_error(NEMO_HIP_ERR_UNSUPPORTED,
                         "%s: ln_bt_av = F (the barotropic *_e levels of ts_rst) is not supported", who);
    for (const auto &v : rst_vars(r, reading))
        if (!v.ptr && !v.opt) return set_error(NEMO_HIP_ERR_ARG, "%s: %s is NULL", who, v.name);
    return 0;
}

// the value after "key": in the manifest text (numbers only)
bool json_num(const std::string &t, const char *key, double *out)
{
    const std::string k = std::string("\"") + key + "\":";
    const size_t q = t.find(k);
    if (q == std::string::npos) return false;
    const char *b = t.c_str() + q + k.size();
    char *e = nullptr;
    *out = strtod(b, &e);
    return e != b;
}

// the dtype of variable `name` in the manifest, "" when absent
std::string json_dtype(const std::string &t, const char *name)
{
    const std::string k = std::string("\"name\": \"") + name + "\"";
    const size_t q = t.find(k);
    if (q == std::string::npos) return "";
    const size_t end = t.find('}', q);
    const size_t d = t.find("\"dtype\": \"", q);
    if (d == std::string::npos || d > end) return "";
    return t.substr(d + 10, 2);
}

}  // namespace

int nemo_hip_rst_write(const char *dir, int kt, const nemo_hip_rst_t *r)
{
    NEED_DOMAIN();
    NEED_NOT_CAPTURING("nemo_hip_rst_write");
    NEED_PTR(dir); NEED_PTR(r);
    int rc = rst_check(r, "rst_write", false);
    if (rc) return rc;
    if (mkdir(dir, 0755) != 0 && errno != EEXIST)
        return set_error(NEMO_HIP_ERR_IO, "rst_write: cannot create %s: %s", dir, strerror(errno));
    HIPCHK(hipStreamSynchronize(g_stream));   // the fields' producers have finished
    const DevDomain &d = g_c->dom;
    char head[512];
    snprintf(head, sizeof head,
             "{\n  \"format\": \"nemo-hip-restart\",\n  \"version\": 1,\n  \"kt\": %d,\n  \"rdt\": %.17g,\n"
             "  \"jpi\": %d,\n  \"jpj\": %d,\n  \"jpk\": %d,\n  \"jperio\": %d,\n  \"variables\": [\n",
             kt, g_c->hdom.rdt, d.jpi, d.jpj, d.jpk, g_c->hdom.jperio);
    std::string man = head;
    std::vector<char> buf;
    std::vector<RstVar> vars;
    for (const auto &v : rst_vars(r))
        if (v.ptr) vars.push_back(v);
    for (size_t q = 0; q < vars.size(); q++) {
        const RstVar &v = vars[q];
        const size_t bytes = v.n * (size_t)v.width;
        const void *src = v.ptr;
        if (g_ptr_mode == NEMO_HIP_PTR_DEVICE) {
            buf.resize(bytes);
            HIPCHK(hipMemcpy(buf.data(), v.ptr, bytes, hipMemcpyDeviceToHost));
            src = buf.data();
        }
        const std::string path = std::string(dir) + "/" + v.name + ".bin";
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) return set_error(NEMO_HIP_ERR_IO, "rst_write: cannot open %s: %s", path.c_str(), strerror(errno));
        const size_t wrote = fwrite(src, 1, bytes, f);
        if (fclose(f) != 0 || wrote != bytes)
            return set_error(NEMO_HIP_ERR_IO, "rst_write: writing %s failed", path.c_str());
        char ent[256];
        if (v.n == d.n3)
            snprintf(ent, sizeof ent, "    {\"name\": \"%s\", \"file\": \"%s.bin\", \"dtype\": \"f%d\", \"shape\": [%d, %d, %d]}",
                     v.name, v.name, v.width, d.jpi, d.jpj, d.jpk);
        else
            snprintf(ent, sizeof ent, "    {\"name\": \"%s\", \"file\": \"%s.bin\", \"dtype\": \"f%d\", \"shape\": [%d, %d]}",
                     v.name, v.name, v.width, d.jpi, d.jpj);
        man += ent;
        man += (q + 1 < vars.size()) ? ",\n" : "\n";
    }
    man += "  ]\n}\n";
    // the manifest last, by rename: a directory with a manifest is complete
    const std::string mp = std::string(dir) + "/manifest.json", tmp = mp + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return set_error(NEMO_HIP_ERR_IO, "rst_write: cannot open %s: %s", tmp.c_str(), strerror(errno));
    const size_t wrote = fwrite(man.data(), 1, man.size(), f);
    if (fclose(f) != 0 || wrote != man.size() || rename(tmp.c_str(), mp.c_str()) != 0)
        return set_error(NEMO_HIP_ERR_IO, "rst_write: writing %s failed", mp.c_str());
    return NEMO_HIP_OK;
}

int nemo_hip_rst_read(const char *dir, nemo_hip_rst_t *r, int *kt, int *neuler)
{
    NEED_DOMAIN();
    NEED_NOT_CAPTURING("nemo_hip_rst_read");
    NEED_PTR(dir); NEED_PTR(r); NEED_PTR(kt); NEED_PTR(neuler);
    int rc = rst_check(r, "rst_read", true);
    if (rc) return rc;
    if (r->un_b && (!r->vn_b || !r->ub_b || !r->vb_b || !r->r1_hu_n || !r->r1_hv_n))
        return set_error(NEMO_HIP_ERR_ARG, "rst_read: un_b given without vn_b, ub_b, vb_b, r1_hu_n and r1_hv_n");
    const std::string mp = std::string(dir) + "/manifest.json";
    std::string man;
    {
        FILE *f = fopen(mp.c_str(), "rb");
        if (!f) return set_error(NEMO_HIP_ERR_IO, "rst_read: cannot open %s: %s", mp.c_str(), strerror(errno));
        char c[4096];
        size_t got;
        while ((got = fread(c, 1, sizeof c, f)) > 0) man.append(c, got);
        fclose(f);
    }
    const DevDomain &d = g_c->dom;
    double v_kt, v_rdt, v_i, v_j, v_k;
    if (man.find("\"format\": \"nemo-hip-restart\"") == std::string::npos || !json_num(man, "kt", &v_kt) ||
        !json_num(man, "rdt", &v_rdt) || !json_num(man, "jpi", &v_i) || !json_num(man, "jpj", &v_j) ||
        !json_num(man, "jpk", &v_k))
        return set_error(NEMO_HIP_ERR_IO, "rst_read: %s is not a restart manifest", mp.c_str());
    if ((int)v_i != d.jpi || (int)v_j != d.jpj || (int)v_k != d.jpk)
        return set_error(NEMO_HIP_ERR_IO, "rst_read: %s holds a %dx%dx%d domain, %dx%dx%d is registered", dir,
                         (int)v_i, (int)v_j, (int)v_k, d.jpi, d.jpj, d.jpk);
    // restart.F90:255-258 and 273-282: a changed time step or no before
    // fields force an Euler restart
    int ne = (v_rdt == g_c->hdom.rdt) ? 1 : 0;
    const bool have_before = !json_dtype(man, "ub").empty();
    if (!have_before) ne = 0;
    const bool read_ts = ne != 0 && d.ln_bt_fw;    // dynspg_ts.F90:1402
    HIPCHK(hipStreamSynchronize(g_stream));          // no queued work still uses the destinations
    const bool dev = (g_ptr_mode == NEMO_HIP_PTR_DEVICE);
    std::vector<char> raw, conv;
    r->kqsr_b = 0;
    for (const auto &v : rst_vars(r, true)) {
        if ((v.before && !have_before) || (v.ts && !read_ts) || !v.ptr) continue;
        const std::string dt = json_dtype(man, v.name);
        if (dt.empty()) {
            // optional: rhop (restart.F90:288-292), the solar heat content
            // and fraction (traqsr.F90:136, 424: absent means an Euler start)
            if (!strcmp(v.name, "rhop") || v.opt) continue;
            return set_error(NEMO_HIP_ERR_IO, "rst_read: %s has no variable %s", dir, v.name);
        }
        const int fw = dt == "f8" ? 8 : dt == "f4" ? 4 : 0;
        if (!fw) return set_error(NEMO_HIP_ERR_IO, "rst_read: %s has dtype %s", v.name, dt.c_str());
        const std::string path = std::string(dir) + "/" + v.name + ".bin";
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) return set_error(NEMO_HIP_ERR_IO, "rst_read: cannot open %s: %s", path.c_str(), strerror(errno));
        raw.resize(v.n * (size_t)fw);
        const size_t got = fread(raw.data(), 1, raw.size(), f);
        const bool extra = fgetc(f) != EOF;
        fclose(f);
        if (got != raw.size() || extra)
            return set_error(NEMO_HIP_ERR_IO, "rst_read: %s does not hold %zu values of %d bytes", path.c_str(), v.n, fw);
        const char *src = raw.data();
        if (fw != v.width) {   // widen f4 to double, or round f8 to float
            conv.resize(v.n * (size_t)v.width);
            for (size_t q = 0; q < v.n; q++) {
                if (v.width == 8) ((double *)conv.data())[q] = (double)((const float *)raw.data())[q];
                else ((float *)conv.data())[q] = (float)((const double *)raw.data())[q];
            }
            src = conv.data();
        }
        const size_t bytes = v.n * (size_t)v.width;
        if (dev) HIPCHK(hipMemcpy(v.ptr, src, bytes, hipMemcpyHostToDevice));
        else memcpy(v.ptr, src, bytes);
        if (!strcmp(v.name, "qsr_hc_b")) r->kqsr_b = 1;
    }
    const size_t n2 = d.n2, n3 = d.n3, w = (size_t)r->real_bytes;
    auto copy = [&](void *dst, const void *src, size_t bytes) -> int {
        if (dev) HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
        else memcpy(dst, src, bytes);
        return 0;
    };
    if (ne == 0) {   // restart.F90:296-307: before fields = now fields
        if ((rc = copy(r->tsb, r->tsn, 2 * n3 * w)) || (rc = copy(r->ub, r->un, n3 * w)) ||
            (rc = copy(r->vb, r->vn, n3 * w)) || (rc = copy(r->sshb, r->sshn, n2 * 8)))
            return rc;
    }
    if (!read_ts) {  // dynspg_ts.F90:1422-1424: start from rest
        for (double *p : {r->ub2_b, r->vb2_b, r->un_bf, r->vn_bf}) {
            if (dev) HIPCHK(hipMemset(p, 0, n2 * 8));
            else memset(p, 0, n2 * 8);
        }
    }
    if (r->un_b) {   // istate.F90:159-178
        Stager S;
        if (r->real_bytes == 8) {
            const double *un = S.in((const double *)r->un, n3), *vn = S.in((const double *)r->vn, n3);
            const double *ub = S.in((const double *)r->ub, n3), *vb = S.in((const double *)r->vb, n3);
            const double *hu = S.in(r->r1_hu_n, n2), *hv = S.in(r->r1_hv_n, n2);
            double *a = S.out(r->un_b, n2), *b = S.out(r->vn_b, n2), *c = S.out(r->ub_b, n2), *e = S.out(r->vb_b, n2);
            rc = S.status ? S.status : launch_istate_bt<double>(un, vn, ub, vb, hu, hv, a, b, c, e);
        } else {
            const float *un = S.in((const float *)r->un, n3), *vn = S.in((const float *)r->vn, n3);
            const float *ub = S.in((const float *)r->ub, n3), *vb = S.in((const float *)r->vb, n3);
            const double *hu = S.in(r->r1_hu_n, n2), *hv = S.in(r->r1_hv_n, n2);
            double *a = S.out(r->un_b, n2), *b = S.out(r->vn_b, n2), *c = S.out(r->ub_b, n2), *e = S.out(r->vb_b, n2);
            rc = S.status ? S.status : launch_istate_bt<float>(un, vn, ub, vb, hu, hv, a, b, c, e);
        }
        if ((rc = S.finish(rc))) return rc;
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    *kt = (int)v_kt;
    *neuler = ne;
    return NEMO_HIP_OK;
}

// ---------------------------------------------------------------------------
// tangent linear and adjoint of the tracer step (one tracer, FP64)
// ---------------------------------------------------------------------------
#define NEED_FCT_ORDERS(h, v)                                                                            \
    do {                                                                                                 \
        if (((h) != 2 && (h) != 4) || ((v) != 2 && (v) != 4))                                           \
            return set_error(NEMO_HIP_ERR_ARG, "kn_fct_h = %d, kn_fct_v = %d (2 or 4)", (h), (v));      \
    } while (0)

int nemo_hip_tra_adv_fct_coef(double p2dt, const double *pun, const double *pvn, const double *pwn, const double *ptb,
                              const double *ptn, int kn_fct_h, int kn_fct_v, double *pcx, double *pcy, double *pcz)
{
    NEED_DOMAIN();
    NEED_FCT_ORDERS(kn_fct_h, kn_fct_v);
    NEED_PTR(pun); NEED_PTR(pvn); NEED_PTR(pwn); NEED_PTR(ptb); NEED_PTR(ptn); NEED_PTR(pcx); NEED_PTR(pcy);
    NEED_PTR(pcz);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const double *u = S.in(pun, n3), *v = S.in(pvn, n3), *w = S.in(pwn, n3), *tb = S.in(ptb, n3), *tn = S.in(ptn, n3);
    double *cx = S.inout(pcx, n3), *cy = S.inout(pcy, n3), *cz = S.inout(pcz, n3);
    int rc = S.status ? S.status : launch_fct_frozen(0, p2dt, u, v, w, tb, tn, kn_fct_h, kn_fct_v, cx, cy, cz, nullptr);
    return S.finish(rc);
}

int nemo_hip_tra_adv_fct_tl(double p2dt, const double *pun, const double *pvn, const double *pwn, const double *ptb,
                            const double *ptn, int kn_fct_h, int kn_fct_v, const double *pcx, const double *pcy,
                            const double *pcz, double *pta)
{
    NEED_DOMAIN();
    NEED_FCT_ORDERS(kn_fct_h, kn_fct_v);
    NEED_PTR(pun); NEED_PTR(pvn); NEED_PTR(pwn); NEED_PTR(ptb); NEED_PTR(ptn); NEED_PTR(pcx); NEED_PTR(pcy);
    NEED_PTR(pcz); NEED_PTR(pta);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const double *u = S.in(pun, n3), *v = S.in(pvn, n3), *w = S.in(pwn, n3), *tb = S.in(ptb, n3), *tn = S.in(ptn, n3);
    const double *cx = S.in(pcx, n3), *cy = S.in(pcy, n3), *cz = S.in(pcz, n3);
    double *ta = S.inout(pta, n3);
    int rc = S.status ? S.status
                      : launch_fct_frozen(1, p2dt, u, v, w, tb, tn, kn_fct_h, kn_fct_v, (double *)cx, (double *)cy,
                                          (double *)cz, ta);
    return S.finish(rc);
}

int nemo_hip_tra_adv_fct_ad(const double *pun, const double *pvn, const double *pwn, int kn_fct_h, int kn_fct_v,
                            const double *pcx, const double *pcy, const double *pcz, const double *pta_ad,
                            double *ptb_ad, double *ptn_ad)
{
    NEED_DOMAIN();
    NEED_FCT_ORDERS(kn_fct_h, kn_fct_v);
    NEED_PTR(pun); NEED_PTR(pvn); NEED_PTR(pwn); NEED_PTR(pcx); NEED_PTR(pcy); NEED_PTR(pcz); NEED_PTR(pta_ad);
    NEED_PTR(ptb_ad); NEED_PTR(ptn_ad);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const double *u = S.in(pun, n3), *v = S.in(pvn, n3), *w = S.in(pwn, n3);
    const double *cx = S.in(pcx, n3), *cy = S.in(pcy, n3), *cz = S.in(pcz, n3), *mu = S.in(pta_ad, n3);
    double *gb = S.inout(ptb_ad, n3), *gn = S.inout(ptn_ad, n3);
    int rc = S.status ? S.status : launch_fct_ad(u, v, w, kn_fct_h, kn_fct_v, cx, cy, cz, mu, gb, gn);
    return S.finish(rc);
}

int nemo_hip_tra_ldf_lap_ad(const double *pahu, const double *pahv, const double *pta_ad, double *ptb_ad)
{
    NEED_DOMAIN();
    NEED_PTR(pahu); NEED_PTR(pahv); NEED_PTR(pta_ad); NEED_PTR(ptb_ad);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const double *a = S.in(pahu, n3), *b = S.in(pahv, n3), *mu = S.in(pta_ad, n3);
    double *gb = S.inout(ptb_ad, n3);
    int rc = S.status ? S.status : launch_ldf_lap_ad(a, b, mu, gb);
    return S.finish(rc);
}

int nemo_hip_tra_zdf_imp_ad(double p2dt, const double *avt, double *pta_ad, double *ptb_ad)
{
    NEED_DOMAIN();
    NEED_PTR(avt); NEED_PTR(pta_ad); NEED_PTR(ptb_ad);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    const double *k = S.in(avt, n3);
    double *ga = S.inout(pta_ad, n3), *gb = S.inout(ptb_ad, n3);
    int rc = S.status ? S.status : launch_zdf_imp_ad(p2dt, k, ga, gb);
    return S.finish(rc);
}

int nemo_hip_tra_nxt_ad(double *ptb_ad, double *ptn_ad, double *pta_ad)
{
    NEED_DOMAIN();
    NEED_PTR(ptb_ad); NEED_PTR(ptn_ad); NEED_PTR(pta_ad);
    const size_t n3 = g_c->dom.n3;
    Stager S;
    double *gb = S.inout(ptb_ad, n3), *gn = S.inout(ptn_ad, n3), *ga = S.out(pta_ad, n3);
    int rc = S.status ? S.status : launch_tra_nxt_ad(gb, gn, ga);
    return S.finish(rc);
}

int nemo_hip_lbc_adj(double *p, int nk)
{
    NEED_DOMAIN();
    NEED_PTR(p);
    if (nk < 1 || nk > g_c->dom.jpk) return set_error(NEMO_HIP_ERR_ARG, "lbc_adj: nk = %d", nk);
    Stager S;
    double *f = S.inout(p, g_c->dom.n2 * (size_t)nk);
    int rc = S.status ? S.status : launch_lbc_adj(f, nk);
    return S.finish(rc);
}

int nemo_hip_eos_register(const nemo_hip_eos_t *eos)
{
    NEED_DOMAIN();
    NEED_NOT_CAPTURING("nemo_hip_eos_register");
    NEED_PTR(eos);
    if (eos->neos != NEMO_NP_SEOS && eos->neos != NEMO_NP_TEOS10 && eos->neos != NEMO_NP_EOS80)
        return set_error(NEMO_HIP_ERR_ARG, "eos_register: neos = %d (TEOS-10 %d, EOS-80 %d or S-EOS %d)", eos->neos,
                         NEMO_NP_TEOS10, NEMO_NP_EOS80, NEMO_NP_SEOS);
    g_c->neos = eos->neos;
    g_c->eos = SeosPar{eos->rn_a0, eos->rn_b0, eos->rn_lambda1, eos->rn_lambda2, eos->rn_mu1, eos->rn_mu2, eos->rn_nu};
    g_c->have_eos = true;
    return NEMO_HIP_OK;
}
int nemo_hip_eos_insitu(const double *pts, double *prd, const double *pdep)
{
    return do_eos<double>(0, pts, prd, nullptr, pdep);
}
int nemo_hip_eos_insitu_pot(const double *pts, double *prd, double *prhop, const double *pdep)
{
    return do_eos<double>(1, pts, prd, prhop, pdep);
}
int nemo_hip_zdf_mxl(int kt, const double *rn2b, const double *gdept, const double *gdepw, int nlb10, double rho_c,
                     int32_t *nmln, double *hmlp, double *hmlpt)
{
    return do_zdf_mxl<double>(kt, rn2b, gdept, gdepw, nlb10, rho_c, nmln, hmlp, hmlpt);
}
int nemo_hip_ldf_slp(int kt, const double *prd, const double *pn2, const double *gdept, const double *gdepw,
                     const int32_t *nmln, const double *hmlp, const double *hmlpt, double rn_slpmax, double *uslp,
                     double *vslp, double *wslpi, double *wslpj)
{
    return do_ldf_slp<double>(kt, prd, pn2, gdept, gdepw, nmln, hmlp, hmlpt, rn_slpmax, uslp, vslp, wslpi, wslpj);
}
int nemo_hip_rab_3d(const double *pts, double *pab, const double *gdept) { return do_rab<double>(pts, pab, gdept); }
int nemo_hip_bn2(const double *pts, const double *pab, double *pn2, const double *gdept, const double *gdepw)
{
    return do_bn2<double>(pts, pab, pn2, gdept, gdepw);
}
int nemo_hip_hpg_zco(int kt, const double *rhd, double *ua, double *va) { return do_hpg_zco<double>(kt, rhd, ua, va); }
int nemo_hip_dyn_ldf_lap(int kt, const double *pub, const double *pvb, double *pua, double *pva, int kpass,
                         const double *ahmt, const double *ahmf)
{
    return do_dyn_ldf_lap<double>(kt, pub, pvb, pua, pva, kpass, ahmt, ahmf);
}

int nemo_hip_stp_ctl(int kt, const double *sshn, const double *un, const double *tsn, double zmax[6],
                     int32_t loc[11], int *kindic)
{
    return do_stp_ctl<double>(kt, sshn, un, tsn, zmax, loc, kindic);
}

int nemo_hip_stream_copy(const void *src, void *dst, size_t bytes, int blocks)
{
    if (!g_init) return set_error(NEMO_HIP_ERR_STATE, "nemo_hip_init has not been called");
    if (!src || !dst) return set_error(NEMO_HIP_ERR_ARG, "stream_copy: NULL pointer");
    if (bytes % 16 || ((uintptr_t)src | (uintptr_t)dst) % 16)
        return set_error(NEMO_HIP_ERR_ARG, "stream_copy: sizes and addresses must be multiples of 16 bytes");
    if (blocks < 1 || blocks > 65536) return set_error(NEMO_HIP_ERR_ARG, "stream_copy: blocks = %d", blocks);
    return launch_stream_copy(src, dst, bytes, blocks);
}

int nemo_hip_tune_set(const char *key, int value)
{
    if (!key) return set_error(NEMO_HIP_ERR_ARG, "tune_set: NULL key");
    for (int q = 0; q < T_NTUNE; q++)
        if (!strcmp(key, g_tune_names[q])) {
            if (value < 0 || value > g_tune_max[q])
                return set_error(NEMO_HIP_ERR_ARG, "tune_set: %s = %d (0 to %d)", key, value, g_tune_max[q]);
            g_tune[q] = value;
            return NEMO_HIP_OK;
        }
    return set_error(NEMO_HIP_ERR_ARG, "tune_set: unknown key %s", key);
}
int nemo_hip_tune_get(const char *key, int *value)
{
    if (!key || !value) return set_error(NEMO_HIP_ERR_ARG, "tune_get: NULL argument");
    for (int q = 0; q < T_NTUNE; q++)
        if (!strcmp(key, g_tune_names[q])) {
            *value = g_tune[q];
            return NEMO_HIP_OK;
        }
    return set_error(NEMO_HIP_ERR_ARG, "tune_get: unknown key %s", key);
}

int nemo_hip_tra_adv(int kt, const double *un, const double *vn, const double *wn, const double *ptb,
                     const double *ptn, double *pta, int kjpt, int kn_fct_h, int kn_fct_v)
{
    return do_tra_adv<double>(kt, un, vn, wn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v);
}
int nemo_hip_tra_adv_sp(int kt, const float *un, const float *vn, const float *wn, const float *ptb,
                        const float *ptn, float *pta, int kjpt, int kn_fct_h, int kn_fct_v)
{
    return do_tra_adv<float>(kt, un, vn, wn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v);
}

int nemo_hip_glob_sum(const double *ptab, int nk, int kind, double *psum)
{
    return do_glob_sum<double>(ptab, nk, kind, psum);
}
int nemo_hip_glob_sum_sp(const float *ptab, int nk, int kind, double *psum)
{
    return do_glob_sum<float>(ptab, nk, kind, psum);
}
int nemo_hip_glob_sum_noalloc(const double *ptab, int nk, int kind, double *psum)
{
    NEED_DOMAIN();
    NEED_PTR(ptab); NEED_PTR(psum);
    if (g_ptr_mode != NEMO_HIP_PTR_DEVICE)
        return set_error(NEMO_HIP_ERR_STATE, "glob_sum_noalloc: device-pointer mode only (host mode would stage)");
    const DevDomain &d = g_c->dom;
    if (nk < 1 || nk > d.jpk) return set_error(NEMO_HIP_ERR_ARG, "glob_sum_noalloc: nk = %d outside 1..jpk", nk);
    if (kind != 0 && kind != 1)
        return set_error(NEMO_HIP_ERR_ARG, "glob_sum_noalloc: kind = %d (0 tmask_i, 1 tmask_h)", kind);
    if (d.n2 * (size_t)nk > (size_t)NEMO_HIP_GLOB_SUM_NOALLOC_MAX)
        return set_error(NEMO_HIP_ERR_ARG, "glob_sum_noalloc: %zu values, more than the maximum %d", d.n2 * (size_t)nk,
                         NEMO_HIP_GLOB_SUM_NOALLOC_MAX);
    double *scr = glob_sum_static_scratch(g_device);
    if (!scr) return set_error(NEMO_HIP_ERR_STATE, "glob_sum_noalloc: no scratch address for device %d", g_device);
    return launch_glob_sum<double>(ptab, nk, kind, scr, psum);
}
int nemo_hip_dia_hsb_init(nemo_hip_hsb_t *h)
{
    if (!h) return set_error(NEMO_HIP_ERR_ARG, "dia_hsb_init: NULL argument");
    if (h->real_bytes == 8) return do_dia_hsb_init<double>(h);
    if (h->real_bytes == 4) return do_dia_hsb_init<float>(h);
    return set_error(NEMO_HIP_ERR_ARG, "dia_hsb_init: real_bytes = %d (8 or 4)", h->real_bytes);
}
int nemo_hip_dia_hsb(int kt, nemo_hip_hsb_t *h)
{
    if (!h) return set_error(NEMO_HIP_ERR_ARG, "dia_hsb: NULL argument");
    if (h->real_bytes == 8) return do_dia_hsb<double>(kt, h);
    if (h->real_bytes == 4) return do_dia_hsb<float>(kt, h);
    return set_error(NEMO_HIP_ERR_ARG, "dia_hsb: real_bytes = %d (8 or 4)", h->real_bytes);
}

int nemo_hip_tra_qsr(int kt, int kqsr_b, int nksr, double rn_abs, double rn_si0, double rn_si1, const double *qsr,
                     const double *gdepw, double *qsr_hc, double *qsr_hc_b, double *tsa, double *fraqsr_1lev)
{
    return do_tra_qsr<double>(kt, kqsr_b, nksr, rn_abs, rn_si0, rn_si1, qsr, gdepw, qsr_hc, qsr_hc_b, tsa,
                              fraqsr_1lev);
}

int nemo_hip_tra_sbc(int kt, double *qns, double *qsr, const double *sfx, const double *emp, const double *tsn,
                     double *tsa, double *sbc_tsc, double *sbc_tsc_b)
{
    return do_tra_sbc<double>(kt, qns, qsr, sfx, emp, tsn, tsa, sbc_tsc, sbc_tsc_b);
}

int nemo_hip_ts_wgt(int ll_av, int ll_fw, int *jpit, double *zwgt1, double *zwgt2)
{
    NEED_DOMAIN();
    NEED_PTR(jpit); NEED_PTR(zwgt1); NEED_PTR(zwgt2);
    return spg_ts_weights(g_c->dom, ll_av, ll_fw, jpit, zwgt1, zwgt2);
}

// ---------------------------------------------------------------------------
// FP32-storage entry points
// ---------------------------------------------------------------------------
int nemo_hip_tra_adv_trp_sp(const float *un, const float *vn, const float *wn, float *zun, float *zvn, float *zwn)
{
    return do_tra_adv_trp<float>(un, vn, wn, zun, zvn, zwn);
}
int nemo_hip_tra_adv_fct_sp(int kt, int kit000, const char *cdtype, double p2dt, const float *pun, const float *pvn,
                            const float *pwn, const float *ptb, const float *ptn, float *pta, int kjpt, int kn_fct_h,
                            int kn_fct_v)
{
    return do_tra_adv_fct<float>(kt, kit000, cdtype, p2dt, pun, pvn, pwn, ptb, ptn, pta, kjpt, kn_fct_h, kn_fct_v);
}
int nemo_hip_tra_ldf_lap_sp(int kt, int kit000, const char *cdtype, const float *pahu, const float *pahv,
                            const float *ptb, float *pta, int kjpt, int kpass)
{
    return do_tra_ldf_lap<float>(kt, kit000, cdtype, pahu, pahv, ptb, pta, kjpt, kpass);
}
int nemo_hip_tra_ldf_iso_sp(int kt, int kit000, const char *cdtype, const float *pahu, const float *pahv,
                            const float *ptb, const float *ptbb, float *pta, int kjpt, int kpass, const float *uslp,
                            const float *vslp, const float *wslpi, const float *wslpj, float *ah_wslp2, float *akz)
{
    return do_tra_ldf_iso<float>(kt, kit000, cdtype, pahu, pahv, ptb, ptbb, pta, kjpt, kpass, uslp, vslp, wslpi, wslpj,
                                 ah_wslp2, akz);
}
int nemo_hip_tra_zdf_sp(int kt, const float *tsb, float *tsa, const float *avt, const float *avs, int l_ldfslp,
                        const float *ah_wslp2, const float *akz)
{
    NEED_DOMAIN();
    const double r2dt = r2dt_update(g_c->r2dt_zdf, kt);
    return do_tra_zdf_imp<float>(kt, g_c->hdom.nit000, "TRA", r2dt, tsb, tsa, 2, avt, avs, l_ldfslp, ah_wslp2, akz, 1);
}
int nemo_hip_dyn_adv_sp(int kt, const float *un, const float *vn, const float *wn, float *ua, float *va)
{
    (void)kt;
    return dyn_adv_dispatch<float>(un, vn, wn, ua, va);
}
int nemo_hip_dyn_vor_sp(int kt, const float *un, const float *vn, float *ua, float *va)
{
    (void)kt;
    return do_dyn_vor<float>(un, vn, ua, va);
}
int nemo_hip_dyn_zdf_sp(int kt, const float *ub, const float *vb, float *ua, float *va, const float *avm,
                        const double *rCdU_bot, const double *ua_b, const double *va_b, const double *utau,
                        const double *utau_b, const double *vtau, const double *vtau_b)
{
    return do_dyn_zdf<float>(kt, ub, vb, ua, va, avm, rCdU_bot, ua_b, va_b, utau, utau_b, vtau, vtau_b);
}
int nemo_hip_dyn_spg_ts_sp(int kt, nemo_hip_spg_ts_t *st) { return do_dyn_spg_ts<float>(kt, st); }

int nemo_hip_step_dyn_sp(int kt, const float *ub, const float *vb, const float *wn, const float *avm,
                         const double *utau_b, const double *vtau_b, nemo_hip_spg_ts_t *st)
{
    return do_step_dyn<float>(kt, ub, vb, wn, avm, utau_b, vtau_b, st);
}

int nemo_hip_step_tra_sp(int kt, const float *un, const float *vn, const float *wn, const float *tsb,
                         const float *tsn, float *tsa, int kn_fct_h, int kn_fct_v, const float *ahtu,
                         const float *ahtv, const float *uslp, const float *vslp, const float *wslpi,
                         const float *wslpj, float *ah_wslp2, float *akz, const float *avt)
{
    return do_step_tra<float>(kt, un, vn, wn, tsb, tsn, tsa, kn_fct_h, kn_fct_v, ahtu, ahtv, uslp, vslp, wslpi,
                              wslpj, ah_wslp2, akz, avt);
}

int nemo_hip_div_hor_sp(int kt, const float *un, const float *vn, float *hdivn)
{
    return do_div_hor<float>(kt, un, vn, hdivn);
}
int nemo_hip_ssh_nxt_sp(int kt, const float *un, const float *vn, float *hdivn, const double *sshb,
                        const double *emp_b, const double *emp, double *ssha)
{
    return do_ssh_nxt<float>(kt, un, vn, hdivn, sshb, emp_b, emp, ssha);
}
int nemo_hip_wzv_sp(int kt, const float *hdivn, float *wn) { return do_wzv<float>(kt, hdivn, wn); }
int nemo_hip_tra_nxt_sp(int kt, float *tsb, float *tsn, float *tsa) { return do_tra_nxt<float>(kt, tsb, tsn, tsa); }
int nemo_hip_dyn_nxt_sp(int kt, float *ub, float *vb, float *un, float *vn, float *ua, float *va,
                        const double *ua_b, const double *va_b, const double *un_adv, const double *vn_adv,
                        double *ub_b, double *vb_b, double *un_b, double *vn_b, const double *r1_hu_n,
                        const double *r1_hv_n)
{
    return do_dyn_nxt<float>(kt, ub, vb, un, vn, ua, va, ua_b, va_b, un_adv, vn_adv, ub_b, vb_b, un_b, vn_b,
                             r1_hu_n, r1_hv_n);
}

int nemo_hip_eos_insitu_sp(const float *pts, float *prd, const float *pdep)
{
    return do_eos<float>(0, pts, prd, nullptr, pdep);
}
int nemo_hip_eos_insitu_pot_sp(const float *pts, float *prd, float *prhop, const float *pdep)
{
    return do_eos<float>(1, pts, prd, prhop, pdep);
}
int nemo_hip_zdf_mxl_sp(int kt, const float *rn2b, const float *gdept, const float *gdepw, int nlb10, double rho_c,
                        int32_t *nmln, double *hmlp, double *hmlpt)
{
    return do_zdf_mxl<float>(kt, rn2b, gdept, gdepw, nlb10, rho_c, nmln, hmlp, hmlpt);
}
int nemo_hip_ldf_slp_sp(int kt, const float *prd, const float *pn2, const float *gdept, const float *gdepw,
                        const int32_t *nmln, const double *hmlp, const double *hmlpt, double rn_slpmax, float *uslp,
                        float *vslp, float *wslpi, float *wslpj)
{
    return do_ldf_slp<float>(kt, prd, pn2, gdept, gdepw, nmln, hmlp, hmlpt, rn_slpmax, uslp, vslp, wslpi, wslpj);
}
int nemo_hip_rab_3d_sp(const float *pts, float *pab, const float *gdept) { return do_rab<float>(pts, pab, gdept); }
int nemo_hip_bn2_sp(const float *pts, const float *pab, float *pn2, const float *gdept, const float *gdepw)
{
    return do_bn2<float>(pts, pab, pn2, gdept, gdepw);
}
int nemo_hip_hpg_zco_sp(int kt, const float *rhd, float *ua, float *va) { return do_hpg_zco<float>(kt, rhd, ua, va); }
int nemo_hip_dyn_ldf_lap_sp(int kt, const float *pub, const float *pvb, float *pua, float *pva, int kpass,
                            const float *ahmt, const float *ahmf)
{
    return do_dyn_ldf_lap<float>(kt, pub, pvb, pua, pva, kpass, ahmt, ahmf);
}

int nemo_hip_stp_ctl_sp(int kt, const double *sshn, const float *un, const float *tsn, double zmax[6],
                        int32_t loc[11], int *kindic)
{
    return do_stp_ctl<float>(kt, sshn, un, tsn, zmax, loc, kindic);
}

int nemo_hip_tra_qsr_sp(int kt, int kqsr_b, int nksr, double rn_abs, double rn_si0, double rn_si1, const double *qsr,
                        const float *gdepw, double *qsr_hc, double *qsr_hc_b, float *tsa, double *fraqsr_1lev)
{
    return do_tra_qsr<float>(kt, kqsr_b, nksr, rn_abs, rn_si0, rn_si1, qsr, gdepw, qsr_hc, qsr_hc_b, tsa,
                             fraqsr_1lev);
}

int nemo_hip_tra_sbc_sp(int kt, double *qns, double *qsr, const double *sfx, const double *emp, const float *tsn,
                        float *tsa, double *sbc_tsc, double *sbc_tsc_b)
{
    return do_tra_sbc<float>(kt, qns, qsr, sfx, emp, tsn, tsa, sbc_tsc, sbc_tsc_b);
}

}  // extern "C"
