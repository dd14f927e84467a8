// Host side of libgodar_amd.so: context, member management, the step
// launcher and the extern "C" entry points declared in include/godar_gpu.h.
#include "godar_gpu.h"
#include "godar_engine.h"
#include "godar_physics.h"
#include "radix_sort.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace godar {
__global__ void k_begin(DS, int, int);
__global__ void k_ranges(DS);
template <bool M> __global__ void k_cand(DS);
template <bool F> __global__ void k_sup(DS);
template <bool F> __global__ void k_sup_dense(DS);
__global__ void k_scan_tiles(DS);
__global__ void k_scan_total(DS, int);
__global__ void k_scan_offsets(DS);
template <bool M> __global__ void k_cand_dense(DS);
template <bool M> __global__ void k_cand_dense_over(DS);
__global__ void k_comp(DS);
template <bool M> __global__ void k_contact(DS);
template <bool M> __global__ void k_bond(DS);
__global__ void k_stress(DS);
template <bool M> __global__ void k_walls(DS);
template <bool M> __global__ void k_repair(DS, int);
__global__ void k_fixq(DS);
template <bool M> __global__ void k_fixup(DS);
template <bool M, bool R> __global__ void k_final(DS);
__global__ void k_totals(DS, int64_t*);

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Working memory of godar_get_totals_noalloc: the totals it returns (pair
// totals summed by k_totals, the three queue totals, the repair totals),
// one copy on each
// GPU the library runs on, held until the process ends.  Host code never
// passes this name to a kernel or a copy: it uses the device address that
// hipGetSymbolAddress gives for the current device (totals_scratch_of).
constexpr int TOTALS_SCRATCH = 32;   // entries: 256 bytes, GODAR_NTOTALS of them used
static __device__ __align__(256) int64_t scratch[TOTALS_SCRATCH];

// The three queue totals (contact, bond, wall work): each is the sum of its
// per-XCD segment counters.  One wavefront.
__global__ void k_queue_totals(const int32_t* qn, int64_t* out) {
    const int t = threadIdx.x;
    if (t < 3) {
        int64_t sum = 0;
        for (int x = 0; x < XCD_COUNT; ++x) sum += qn[(t * XCD_COUNT + x) * QN_STRIDE];
        out[t] = sum;
    }
}

// Ridging seeds and events re-evaluated by the repair in the last step each
// active member took, summed over the members (one block).
__global__ void k_member_totals(DS S, int mcap, int64_t* out) {
    __shared__ int64_t part[2][256];
    int64_t a = 0, b = 0;
    for (int m = threadIdx.x; m < mcap; m += blockDim.x)
        if (S.members[m].active) { a += S.stats[4 * m + 2]; b += S.stats[4 * m + 3]; }
    part[0][threadIdx.x] = a;
    part[1][threadIdx.x] = b;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            part[0][threadIdx.x] += part[0][threadIdx.x + w];
            part[1][threadIdx.x] += part[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = part[0][0]; out[1] = part[1][0]; }
}

__global__ void k_zero_i64(int64_t* p, int n) {
    const int t = threadIdx.x;
    if (t < n) p[t] = 0;
}
}  // namespace godar

using namespace godar;

struct Range { int64_t first, n; };

struct godar_ctx {
    hipStream_t stream = nullptr;
    // The contact and wall kernels run beside the bond kernel, on a side
    // stream forked from and joined back to `stream` by events (a captured
    // step graph takes the same branches); steps timed stage by stage run
    // them one after another.
    hipStream_t side = nullptr;
    hipEvent_t efork = nullptr, ejoin = nullptr;
    int precision = GODAR_FP64;
    int64_t S = 0;
    int Mcap = 0;
    int CO = 32, CI = 32, HC = 16, BC = 16;
    int SEEDK = 2, HEAPK = 16, TABK = 4;
    int block = 256;
    int sblock[GODAR_NSTAGES];          // threads per block of each stage's per-floe kernels
    int record = 0;
    int parity = 0;
    int32_t N_used = 0;
    int32_t ncells_cap = 0;
    std::vector<Member> members;
    std::vector<DevParams> dparams;
    std::vector<int32_t> cell_n;        // cells owned by each member
    // per member: local index of the floe held by storage slot first + t
    std::vector<std::vector<int32_t>> perm;
    std::vector<WindGridDesc> wgrids;
    std::vector<double> wpool;
    int64_t wpool_dev_cap = 0;
    int wgrids_dev_cap = 0;
    DS d{};
    // double-buffered pair state
    HistRec* hist[2] = {nullptr, nullptr};
    int32_t* hcnt[2] = {nullptr, nullptr};
    BondRec* bond[2] = {nullptr, nullptr};
    int32_t* bcnt[2] = {nullptr, nullptr};
    RadixScratch rs;
    std::string err;
    // stage timing with HIP events on the context stream (godar_profile)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double stage_ms[GODAR_NSTAGES] = {0};
    int64_t stage_n[GODAR_NSTAGES] = {0};
    // state of the last complete step, for rolling back an aborted step
    int pend_parity = 0;
    int64_t pend_done0 = 0;
    bool pending = false;
    uint32_t last_err_bits = 0;
    bool auto_grow = true;
    // captured step graphs, one per pair-state parity (the kernels take the
    // device-state pointers by value), each with the launch signature it was
    // captured for
    bool use_graphs = true;
    int dense_scan = 96;                // floes per 3x3 search above which a member is dense
    int repair_lanes = 64;              // repair events evaluated at once
    int repair_cap = 1024;              // pending repair keys held in LDS
    double skin_frac = 0.1;             // superset skin, x (max r + rtree)
    int thread_span = 1;                // cells searched on each side by the thread search
    double tuned_cost_ms = 0.0;         // the autotuner's price of the chosen search shape
    std::vector<double> rmax;           // largest radius of each member when it was added
    int64_t sup_cap = 0;                // superset pool entries
    int32_t built_qseg = -1;            // queue share the dense list was built with
    // captured step graphs by [parity][lean]
    hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    std::vector<unsigned char> gsig[2][2];
    // Steps without a build launch no build kernels (lean steps): the host
    // launches the full sequence when it forced a build or refresh itself,
    // after LEAN_RUN lean steps in a row (the device builds early there once
    // a floe has moved 0.3 skin), and for a while after a lean step found a
    // build due (ERR_NEED_BUILD; godar_sync runs such steps again).
    bool need_full = true;
    int lean_run = 0;                   // lean steps since the last full one
    int full_hold = 0;                  // full steps still to launch after a missed build
    int64_t pend_launched = 0;          // steps launched since the last sync
    int device = 0;                     // the GPU the context lives on
    int ncu = 256;                      // its compute units
    int64_t* totals_scratch = nullptr;  // device address of that GPU's `scratch`
};

namespace {

int fail(godar_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((c), GODAR_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    return hipMalloc((void**)p, sizeof(T) * count);
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void derive_params(const godar_params& p, DevParams& d) {
    std::memset(&d, 0, sizeof(d));
    d.dt = p.dt; d.rtree = p.rtree; d.nx = p.nx; d.ny = p.ny;
    d.rhoair = p.rhoair; d.rhoice = p.rhoice; d.rhowater = p.rhowater;
    d.Cdair = p.Cdair; d.Cdwater = p.Cdwater; d.Csair = p.Csair; d.Cswater = p.Cswater;
    d.z0w = p.z0w;
    // par_get.f90:35,64-68,263-266
    d.pi = 4.0 * std::atan(1.0);
    d.ec = p.e_modul / (2.0 * (1.0 - p.poiss_ratio * p.poiss_ratio));
    d.gc = p.e_modul / (4.0 * (1.0 - p.poiss_ratio) * (2.0 + p.poiss_ratio));
    double lr = std::log(p.rest_coeff);
    d.beta = 10.0 * lr / std::sqrt(lr * lr + d.pi * d.pi);
    d.pi_ec = d.pi * d.ec;
    d.k6 = 6.0 * d.gc / d.ec;
    d.two_beta = 2.0 * d.beta;
    d.sinlat = std::sin(p.lat);
    d.log10z0 = std::log(10.0 / p.z0w);
    d.friction_coeff = p.friction_coeff; d.sigmanc_crit = p.sigmanc_crit;
    d.eb = p.eb; d.lambda_rb = p.lambda_rb; d.lambda_lb = p.lambda_lb;
    d.sigmatb_crit = p.sigmatb_crit; d.sigmacb_crit = p.sigmacb_crit; d.tau_crit = p.tau_crit;
    d.bond_lim = p.bond_lim; d.dmax = p.dmax; d.dtd = p.dtd; d.dth = p.dth;
    d.ua = p.ua; d.va = p.va; d.uw = p.uw; d.vw = p.vw;
    d.slipping = p.slipping; d.cohesion = p.cohesion; d.ridging = p.ridging;
    d.shelter = p.shelter; d.restart = p.restart;
    double fr = p.bond_frac;
    d.bond_thr24 = fr >= 1.0 ? (1u << 24) : (fr <= 0.0 ? 0u : (uint32_t)std::floor(fr * 16777216.0));
    d.bond_seed = p.bond_seed;
    d.wind_grid = -1;
}

// The per-floe lists sized by the list capacities (CO, CI, HC, BC).
struct PairLists {
    int64_t* oj = nullptr;
    double* od = nullptr;
    int64_t* ik = nullptr;
    uint32_t* evf = nullptr;
    int32_t* ov = nullptr;
    BondTerms* bterm = nullptr;
    double *ssc = nullptr, *ssb = nullptr;
    int32_t* ibond = nullptr;
    HistRec* hist[2] = {nullptr, nullptr};
    BondRec* bond[2] = {nullptr, nullptr};
};

void free_lists(PairLists& L) {
    dfree(L.oj); dfree(L.od); dfree(L.ik); dfree(L.evf); dfree(L.ov); dfree(L.bterm);
    dfree(L.ssc); dfree(L.ssb); dfree(L.ibond);
    for (int k = 0; k < 2; ++k) { dfree(L.hist[k]); dfree(L.bond[k]); }
}

// All or nothing: on failure nothing stays allocated.
hipError_t alloc_lists(PairLists& L, int64_t S, int CO, int CI, int HC, int BC) {
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t x) { e = x; return x == hipSuccess; };
    bool good = ok(dalloc(&L.oj, S * CO)) && ok(dalloc(&L.od, S * CO)) && ok(dalloc(&L.ik, S * CI)) &&
                ok(dalloc(&L.evf, S * CO)) && ok(dalloc(&L.ov, S * CO)) &&
                ok(dalloc(&L.bterm, S * BC)) && ok(dalloc(&L.ssc, S * CO)) &&
                ok(dalloc(&L.ibond, S * CI)) && ok(dalloc(&L.ssb, S * CO));
    for (int k = 0; good && k < 2; ++k)
        good = ok(dalloc(&L.hist[k], S * HC)) && ok(dalloc(&L.bond[k], S * BC));
    if (!good) free_lists(L);
    return e;
}

PairLists lists_of(const godar_ctx* c) {
    const DS& d = c->d;
    PairLists L;
    L.oj = d.oj; L.od = d.od; L.ik = d.ik; L.evf = d.evf; L.ov = d.ov; L.bterm = d.bterm;
    L.ssc = d.ssc; L.ssb = d.ssb; L.ibond = d.ibond;
    for (int k = 0; k < 2; ++k) { L.hist[k] = c->hist[k]; L.bond[k] = c->bond[k]; }
    return L;
}

void set_lists(godar_ctx* c, const PairLists& L) {
    DS& d = c->d;
    d.oj = L.oj; d.od = L.od; d.ik = L.ik; d.evf = L.evf; d.ov = L.ov; d.bterm = L.bterm;
    d.ssc = L.ssc; d.ssb = L.ssb; d.ibond = L.ibond;
    for (int k = 0; k < 2; ++k) { c->hist[k] = L.hist[k]; c->bond[k] = L.bond[k]; }
}

void bind_parity(godar_ctx* c) {
    int o = c->parity, n = 1 - c->parity;
    c->d.hist_old = c->hist[o]; c->d.hist_new = c->hist[n];
    c->d.hcnt_old = c->hcnt[o]; c->d.hcnt_new = c->hcnt[n];
    c->d.bond_old = c->bond[o]; c->d.bond_new = c->bond[n];
    c->d.bcnt_old = c->bcnt[o]; c->d.bcnt_new = c->bcnt[n];
}

// The next step builds the candidate supersets (and the cell order) anew.
void force_rebuild(godar_ctx* c) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->d.rebuild, 1, BUILD_DUE);
    c->need_full = true;
}

// The next step copies the floe records into the current order anew (the
// host changed what the copies hold, e.g. the wind).
void force_refresh(godar_ctx* c) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->d.rebuild + 1, 1, 1);
    c->need_full = true;
}

int32_t qseg_of(const godar_ctx* c) {
    const int bc = c->sblock[3];
    const int gc = ((std::max(c->N_used, 1) + bc - 1) / bc + XCD_COUNT - 1) / XCD_COUNT * XCD_COUNT;
    return gc / XCD_COUNT * bc;
}

void sync_ds(godar_ctx* c) {
    DS& d = c->d;
    d.S = c->S; d.N = c->N_used;
    d.CO = c->CO; d.CI = c->CI; d.HC = c->HC; d.BC = c->BC;
    d.SEEDK = c->SEEDK; d.HEAPK = c->HEAPK; d.TABK = c->TABK;
    d.record_forces = c->record;
    d.ncells = c->ncells_cap;
    d.qseg = qseg_of(c);   // XCD share of the candidate search (k_cand's grid, XCD-aware order)
    d.repair_lanes = c->repair_lanes;
    d.repair_cap = c->repair_cap;
    d.sup_cap = c->sup_cap;
    d.bkey = c->rs.keys_alt;
    d.bval = c->rs.vals_alt;
    bind_parity(c);
}

// Floe records up to date with the ordered copies (a no-op unless a step ran
// since the last spill), before the host reads or writes them or changes the
// floe count; the next step then reads rec as it stands.
void spill(godar_ctx* c) {
    if (c->N_used == 0) return;
    sync_ds(c);
    hipLaunchKernelGGL(k_spill, dim3((c->N_used + 255) / 256), dim3(256), 0, c->stream, c->d, 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->d.rebuild + 2, 1, 0);
}

int ensure_cells(godar_ctx* c, int32_t need) {
    if (need <= c->ncells_cap) return GODAR_OK;
    int32_t cap = std::max(need, c->ncells_cap * 2);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int2* cr = nullptr;
    if (dalloc(&cr, (int64_t)cap + 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, GODAR_ERR_NOMEM, "out of device memory for the cell table");
    }
    dfree(c->d.crange);
    c->d.crange = cr;
    c->ncells_cap = cap;
    return GODAR_OK;
}

int upload_tables(godar_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->d.members, c->members.data(), sizeof(Member) * c->Mcap,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync((void*)c->d.params, c->dparams.data(), sizeof(DevParams) * c->Mcap,
                             hipMemcpyHostToDevice, c->stream));
    return GODAR_OK;
}

int upload_members_only(godar_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->d.members, c->members.data(), sizeof(Member) * c->Mcap,
                             hipMemcpyHostToDevice, c->stream));
    return GODAR_OK;
}

// steps completed without error since the context was created
int steps_done(godar_ctx* c, int64_t* out) {
    HIPCHK(c, hipMemcpyAsync(out, c->d.steps_done, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GODAR_OK;
}

int download_members(godar_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->members.data(), c->d.members, sizeof(Member) * c->Mcap,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GODAR_OK;
}

// smallest free range of length n (first fit over the gaps between members)
bool find_range(const godar_ctx* c, int64_t n, int64_t cap, const std::vector<Range>& used,
                int64_t& first) {
    std::vector<Range> u = used;
    std::sort(u.begin(), u.end(), [](const Range& a, const Range& b) { return a.first < b.first; });
    int64_t pos = 0;
    for (const Range& r : u) {
        if (r.first - pos >= n) { first = pos; return true; }
        pos = std::max(pos, r.first + r.n);
    }
    if (cap - pos >= n) { first = pos; return true; }
    (void)c;
    return false;
}

// Device address of `scratch` on the current GPU, resolved once per GPU
// (each GPU holds its own copy of the array at its own address).
std::mutex g_scratch_mu;
std::vector<int64_t*> g_scratch_addr;

hipError_t totals_scratch_of(int dev, int64_t** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_addr.size() <= dev) g_scratch_addr.resize(dev + 1, nullptr);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[dev] = static_cast<int64_t*>(p);
    }
    *out = g_scratch_addr[dev];
    return hipSuccess;
}

bool valid_member(const godar_ctx* c, int32_t m) {
    return c && m >= 0 && m < c->Mcap && c->members[m].active;
}

hipEvent_t take_event(godar_ctx* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct StageTimer {
    godar_ctx* c;
    int stage;
    hipEvent_t a = nullptr, b = nullptr;
    StageTimer(godar_ctx* c_, int s_) : c(c_), stage(s_) {
        if (!c->prof) return;
        a = take_event(c);
        b = take_event(c);
        if (a) (void)hipEventRecord(a, c->stream);
    }
    ~StageTimer() {
        if (!c->prof || !a || !b) return;
        (void)hipEventRecord(b, c->stream);
        c->ev_pending.push_back({stage, {a, b}});
    }
};

void collect_events(godar_ctx* c) {
    for (auto& p : c->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            c->stage_ms[p.first] += ms;
            c->stage_n[p.first] += 1;
        }
        c->ev_pool.push_back(p.second.first);
        c->ev_pool.push_back(p.second.second);
    }
    c->ev_pending.clear();
}

// grid of a per-floe kernel of stage st over N items; kernels with the
// XCD-aware block order need a multiple of XCD_COUNT blocks
static dim3 grid_for(int N, int B, bool xcd) {
    int g = (N + B - 1) / B;
    if (xcd) g = (g + XCD_COUNT - 1) / XCD_COUNT * XCD_COUNT;
    return dim3(g);
}
// Queue kernels loop over their entries: at most QUEUE_GRID_WAVES
// wavefronts (about one round of resident wavefronts), so that short queues
// do not launch a wavefront per 64 floes that finds nothing to do.
static dim3 grid_queue(int N, int B) {
    const int cap = QUEUE_GRID_WAVES * 64 / B / XCD_COUNT * XCD_COUNT;
    const dim3 g = grid_for(N, B, true);
    return dim3(std::max(XCD_COUNT, std::min((int)g.x, cap)));
}

constexpr int LEAN_RUN = 15;       // lean steps between two full ones
constexpr int FULL_HOLD = 64;      // full steps after a lean step found a build due

int launch_step(godar_ctx* c, bool lean) {
    DS& d = c->d;
    sync_ds(c);
    int N = c->N_used;
    if (N == 0) return GODAR_OK;
    const int* B = c->sblock;
    hipStream_t s = c->stream;
    // Steps that start with d.rebuild set bin, sort and search the cells and
    // build the candidate supersets; the other steps' build kernels return at
    // once (the launch sequence, and a captured graph, is the same).
    {
        StageTimer t(c, 0);
        // (a lean step only resets the queue counters and the members' words)
        const int nt = lean ? std::max(c->Mcap, NQUEUE * XCD_COUNT)
                            : std::max({N, c->ncells_cap + 1, NQUEUE * XCD_COUNT, c->Mcap});
        hipLaunchKernelGGL(k_begin, grid_for(nt, B[0], false), dim3(B[0]), 0, s, d, lean ? 1 : 0,
                           (int)c->Mcap);
    }
    int bits = 1;
    while ((1u << bits) <= (uint32_t)c->ncells_cap && bits < 32) ++bits;
    if (!lean) {
        StageTimer t(c, 1);
        HIPCHK(c, radix_sort_pairs(c->rs, d.ckey, d.cval, N, bits, d.err, d.rebuild, s));
    }
    int nm = c->Mcap;
    const bool mixed = c->precision == GODAR_MIXED;
    bool any_dense = false;
    int64_t dense_n = 0;
    for (int k = 0; k < c->Mcap; ++k)
        if (c->members[k].active && c->dparams[k].dense) { any_dense = true; dense_n += c->members[k].n; }
    // The dense search runs a thread per dense floe (blocks of one
    // wavefront, or of two that split the supersets when the dense floes
    // fit on the GPU at once; each block with its list of near pairs in
    // LDS); the superset
    // builds keep 1,024 wavefronts per XCD: they return at once on most
    // steps, and dispatching many more would cost more than they save.
    const dim3 dense_grid(XCD_COUNT * (int)std::max<int64_t>(1, (dense_n + XCD_COUNT * 64 - 1) / (XCD_COUNT * 64)));
    if (!lean) {
        StageTimer t(c, 2);
        hipLaunchKernelGGL(k_ranges, grid_for(N, B[2], false), dim3(B[2]), 0, s, d);
        // superset build: counts, offsets, entries (k_sup shares k_cand's
        // grid, whose XCD shares lay out the dense-search list)
        const int ntiles = (N + SCAN_TILE - 1) / SCAN_TILE;
        hipLaunchKernelGGL(k_sup<false>, grid_for(N, B[3], true), dim3(B[3]), 0, s, d);
        if (any_dense) hipLaunchKernelGGL(k_sup_dense<false>, dim3(XCD_COUNT * 1024), dim3(64), 0, s, d);
        hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(SCAN_TILE), 0, s, d);
        hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(SCAN_TILE), 0, s, d, ntiles);
        hipLaunchKernelGGL(k_scan_offsets, dim3(ntiles), dim3(SCAN_TILE), 0, s, d);
        hipLaunchKernelGGL(k_sup<true>, grid_for(N, B[3], true), dim3(B[3]), 0, s, d);
        if (any_dense) hipLaunchKernelGGL(k_sup_dense<true>, dim3(XCD_COUNT * 1024), dim3(64), 0, s, d);
    }
    {
        StageTimer t(c, 3);
        // LDS: one list of (floe, candidate) pairs per wavefront
        hipLaunchKernelGGL(mixed ? k_cand<true> : k_cand<false>, grid_for(N, B[3], true),
                           dim3(B[3]), (B[3] / 64) * CAND_LDS_PER_WAVE, s, d);
        if (any_dense) {
            // a thread per dense floe, a contiguous share of the list per XCD
            // two wavefronts per 64 floes (one per half of the supersets)
            // when they all fit on the GPU at once: a shorter chain per
            // wavefront; with more floes, one wavefront takes both halves
            const bool two = dense_n <= (int64_t)c->ncu * 4 * DENSE_WAVES_PER_SIMD * 32;
            auto kd = mixed ? (two ? k_cand_dense<true, true> : k_cand_dense<true, false>)
                            : (two ? k_cand_dense<false, true> : k_cand_dense<false, false>);
            hipLaunchKernelGGL(kd, dense_grid, dim3(two ? 128 : 64), 0, s, d);
            hipLaunchKernelGGL(mixed ? k_cand_dense_over<true> : k_cand_dense_over<false>,
                               grid_for(N, FLOE_BLOCK, true), dim3(FLOE_BLOCK), 0, s, d);
        }
    }
    // contacts, bonds and walls touch disjoint state (the flag words both
    // the contact and the bond kernel set are set with atomicOr), so they may
    // run together: contacts and walls on the side stream, bonds (the
    // longest) on the step's stream, whose next kernel waits for the side
    // stream once
    const bool fork = !c->prof;
    hipStream_t sc = fork ? c->side : s;
    if (fork) {
        (void)hipEventRecord(c->efork, s);
        (void)hipStreamWaitEvent(sc, c->efork, 0);
    }
    {
        StageTimer t(c, 4);
        // LDS: one pass of 64 pair results per wavefront
        hipLaunchKernelGGL(mixed ? k_contact<true> : k_contact<false>, grid_queue(N, B[4]),
                           dim3(B[4]), (B[4] / 64) * CONTACT_LDS_PER_WAVE, sc, d);
    }
    {
        StageTimer t(c, 7);
        hipLaunchKernelGGL(mixed ? k_walls<true> : k_walls<false>, grid_queue(N, B[7]),
                           dim3(B[7]), 0, sc, d);
    }
    {
        StageTimer t(c, 5);
        hipLaunchKernelGGL(mixed ? k_bond<true> : k_bond<false>, grid_queue(N, B[5]),
                           dim3(B[5]), 0, s, d);
    }
    if (fork) {
        (void)hipEventRecord(c->ejoin, sc);
        (void)hipStreamWaitEvent(s, c->ejoin, 0);
    }
    {
        StageTimer t(c, 8);
        // contact components (roots), then the ordered repair
        hipLaunchKernelGGL(k_comp, grid_queue(N, FLOE_BLOCK), dim3(FLOE_BLOCK), 0, s, d);
        hipLaunchKernelGGL(mixed ? k_repair<true> : k_repair<false>, dim3(nm), dim3(64), 0, s, d, nm);
    }
    {
        StageTimer t(c, 9);
        // the touched floes into the fixup queue (laid out as the candidate
        // search's grid, so that a wavefront's positions lie in one XCD
        // share), then a few of them per wavefront
        hipLaunchKernelGGL(k_fixq, grid_for(N, B[3], true), dim3(B[3]), 0, s, d);
        const int fb = std::max(XCD_COUNT, FIXUP_GRID_WAVES * 64 / B[9] / XCD_COUNT * XCD_COUNT);
        hipLaunchKernelGGL(mixed ? k_fixup<true> : k_fixup<false>, dim3(fb),
                           dim3(B[9]), (B[9] / 64) * FIX_LDS_PER_WAVE, s, d);
    }
    if (c->record) {
        StageTimer t(c, 10);
        hipLaunchKernelGGL(k_stress, grid_for(N, B[10], true), dim3(B[10]), 0, s, d);
    }
    {
        StageTimer t(c, 11);
        auto kf = mixed ? (c->record ? k_final<true, true> : k_final<true, false>)
                        : (c->record ? k_final<false, true> : k_final<false, false>);
        hipLaunchKernelGGL(kf, grid_for(N, B[11], false), dim3(B[11]), 0, s, d);
    }
    HIPCHK(c, hipGetLastError());
    c->parity ^= 1;
    return GODAR_OK;
}

// Everything a captured step depends on: the kernels' by-value state and
// the launch shapes (radix scratch pointers and sizes included).
std::vector<unsigned char> launch_signature(godar_ctx* c) {
    sync_ds(c);
    std::vector<unsigned char> sig(sizeof(DS) + sizeof(RadixScratch) + sizeof(c->sblock) +
                                   5 * sizeof(int));
    unsigned char* p = sig.data();
    std::memcpy(p, &c->d, sizeof(DS)); p += sizeof(DS);
    std::memcpy(p, &c->rs, sizeof(RadixScratch)); p += sizeof(RadixScratch);
    std::memcpy(p, c->sblock, sizeof(c->sblock)); p += sizeof(c->sblock);
    // the dense kernels are launched, and their grid sized, by the members'
    // dense floe count
    int dense_members = 0, dense_n = 0;
    for (int k = 0; k < c->Mcap; ++k)
        if (c->members[k].active && c->dparams[k].dense) { ++dense_members; dense_n += c->members[k].n; }
    int extra[5] = {c->precision, c->Mcap, c->N_used, dense_members, dense_n};
    std::memcpy(p, extra, sizeof(extra));
    return sig;
}

void drop_graphs(godar_ctx* c) {
    for (int k = 0; k < 2; ++k)
        for (int l = 0; l < 2; ++l) {
            if (c->gexec[k][l]) (void)hipGraphExecDestroy(c->gexec[k][l]);
            c->gexec[k][l] = nullptr;
            c->gsig[k][l].clear();
        }
}

// One step as a replay of the captured graph of the current parity; the
// graph is (re)captured when the launch signature changed.
int graph_step(godar_ctx* c, bool lean) {
    if (c->N_used == 0) return GODAR_OK;
    const int par = c->parity, l = lean ? 1 : 0;
    std::vector<unsigned char> sig = launch_signature(c);
    if (!c->gexec[par][l] || sig != c->gsig[par][l]) {
        if (c->gexec[par][l]) (void)hipGraphExecDestroy(c->gexec[par][l]);
        c->gexec[par][l] = nullptr;
        c->gsig[par][l].clear();
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = launch_step(c, lean);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        c->parity = par;   // the captured step has not run yet
        bind_parity(c);
        if (rc != GODAR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIPCHK(c, e);
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(c, e);
        c->gexec[par][l] = x;
        c->gsig[par][l] = std::move(sig);
    }
    HIPCHK(c, hipGraphLaunch(c->gexec[par][l], c->stream));
    c->parity ^= 1;
    return GODAR_OK;
}

std::string err_text(uint32_t e) {
    std::string s;
    if (e & ERR_OUTER_CAP) s += " outer-candidate list full;";
    if (e & ERR_INNER_CAP) s += " inner-candidate list full;";
    if (e & ERR_HIST_CAP) s += " contact-history list full;";
    if (e & ERR_BOND_CAP) s += " bond list full;";
    if (e & ERR_REPAIR_CAP) s += " ridging repair table full;";
    if (e & ERR_SEED_CAP) s += " ridging seed list full;";
    if (e & ERR_SUP_CAP) s += " candidate superset pool full;";
    if (e & ERR_NEED_BUILD) s += " build due in a lean step;";
    return s;
}

}  // namespace

// ===================================================================== ABI
extern "C" {

int godar_params_default(godar_params* p) {
    if (!p) return GODAR_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    // par_get.f90:1-111
    p->dt = 1e-3; p->rtree = 1e3; p->nx = 5e3; p->ny = 5e3;
    p->rhoair = 1.3; p->rhoice = 9e2; p->rhowater = 1026.0;
    p->Cdair = 3e-1; p->Cdwater = 3e-1; p->Csair = 5e-4; p->Cswater = 2e-3;
    p->z0w = 3.0e-4;
    p->lat = (double)1.4f;   // default-real literal assigned to a double (par_get.f90:47)
    p->e_modul = 6e9; p->poiss_ratio = 33e-2; p->friction_coeff = 7e-1;
    p->rest_coeff = 88e-2; p->sigmanc_crit = 1e6;
    p->eb = 6e9; p->lambda_rb = 8e-1; p->lambda_lb = 1.0;
    p->sigmatb_crit = 1e5; p->sigmacb_crit = 1e6; p->tau_crit = 1e6;
    p->bond_lim = 1e-2; p->dmax = 9e-1; p->dtd = 1.0; p->dth = 1e5;
    p->bond_frac = 1.0;
    p->ntree = 1; p->slipping = 1; p->cohesion = 1; p->ridging = 1; p->shelter = 1;
    p->restart = 0; p->bond_seed = 0;
    return GODAR_OK;
}

int godar_create(int64_t max_floes, int32_t max_members, int32_t precision, godar_ctx** out) {
    if (!out || max_floes <= 0 || max_floes > (int64_t)1 << 30 || max_members <= 0 ||
        (precision != GODAR_FP64 && precision != GODAR_MIXED))
        return GODAR_ERR_ARG;
    *out = nullptr;
    godar_ctx* c = new godar_ctx();
    c->precision = precision;
    // default block sizes: 256 threads, except the bond and wall stages,
    // which run beside the contact stage (64), and the integration (128);
    // measured on the 512-field ensemble and the served mix
    for (int k = 0; k < GODAR_NSTAGES; ++k) c->sblock[k] = FLOE_BLOCK;
    c->sblock[5] = c->sblock[7] = 64;
    c->sblock[11] = 128;
    c->S = max_floes;
    c->Mcap = max_members;
    c->members.assign(max_members, Member{0, 0, 1, 0, 0});
    c->dparams.assign(max_members, DevParams{});
    c->cell_n.assign(max_members, 0);
    c->rmax.assign(max_members, 0.0);
    c->perm.assign(max_members, {});
    for (auto& m : c->members) m.active = 0;
    DS& d = c->d;
    int64_t S = c->S;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ejoin, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->efork, hipEventDisableTiming);
#define ACHK(x)                                        \
    do {                                               \
        e = (x);                                       \
        if (e != hipSuccess) {                         \
            godar_destroy(c);                          \
            return GODAR_ERR_NOMEM;                    \
        }                                              \
    } while (0)
    if (e != hipSuccess) { godar_destroy(c); return GODAR_ERR_HIP; }
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        totals_scratch_of(c->device, &c->totals_scratch) != hipSuccess) {
        godar_destroy(c);
        return GODAR_ERR_HIP;
    }
    ACHK(dalloc(&d.rec, S));
    ACHK(dalloc(&d.geo, S));
    ACHK(dalloc(&d.shel, S));
    ACHK(dalloc(&d.kin, S));
    ACHK(dalloc(&d.pos_of, S));
    ACHK(dalloc(&d.pos_id, S));
    ACHK(dalloc(&d.wbc, 4 * S));
    ACHK(dalloc(&d.member_of, S));
    ACHK(dalloc(&d.gid_of, S));
    ACHK(dalloc(&d.sid_of, S));
    ACHK(dalloc(&d.members, max_members));
    DevParams* pp;
    ACHK(dalloc(&pp, max_members));
    d.params = pp;
    ACHK(dalloc(&d.ckey, S));
    ACHK(dalloc(&d.cval, S));
    ACHK(dalloc(&d.ocnt, S));
    ACHK(dalloc(&d.icnt, S));
    for (int k = 0; k < 2; ++k) {
        ACHK(dalloc(&c->hcnt[k], S));
        ACHK(dalloc(&c->bcnt[k], S));
    }
    {
        PairLists L;
        ACHK(alloc_lists(L, S, c->CO, c->CI, c->HC, c->BC));
        set_lists(c, L);
    }
    ACHK(dalloc(&d.acc, (int64_t)NACC * S));
    ACHK(dalloc(&d.wout, S));
    ACHK(dalloc(&d.wflag, (S + 3) / 4 * 4));   // whole 32-bit words (wflag_set)
    ACHK(dalloc(&d.cls, S));
    ACHK(dalloc(&d.qlist, (int64_t)NQUEUE * S));
    ACHK(dalloc(&d.bgat, S));
    ACHK(dalloc(&d.qn, (int64_t)NQUEUE * XCD_COUNT * QN_STRIDE));
    ACHK(dalloc(&d.seeds, (int64_t)c->SEEDK * S));
    ACHK(dalloc(&d.nseeds, max_members));
    ACHK(dalloc(&d.heap, (int64_t)c->HEAPK * S));
    ACHK(dalloc(&d.etab, (int64_t)c->TABK * S));
    ACHK(dalloc(&d.netab, max_members));
    ACHK(dalloc(&d.wtab, S));
    ACHK(dalloc(&d.nwtab, max_members));
    ACHK(dalloc(&d.ovw, S));
    ACHK(dalloc(&d.cur, 7 * S));
    ACHK(dalloc(&d.dirty, (S + 3) / 4 * 4));   // whole 32-bit words (dirty_set)
    ACHK(dalloc(&d.comp, S));
    ACHK(dalloc(&d.wnz, S));
    ACHK(dalloc(&d.stats, 4 * (int64_t)max_members));
    ACHK(dalloc(&d.tcur, max_members));
    ACHK(dalloc(&d.fout, 33 * S));
    ACHK(dalloc(&d.err, 1));
    ACHK(dalloc(&d.steps_done, 1));
    c->sup_cap = 8 * S;
    ACHK(dalloc(&d.sup, c->sup_cap));
    ACHK(dalloc(&d.scnt, S));
    ACHK(dalloc(&d.sout, S));
    ACHK(dalloc(&d.soff, S));
    ACHK(dalloc(&d.sblk, S / SCAN_TILE + 2));
    ACHK(dalloc(&d.x0, S));
    ACHK(dalloc(&d.rebuild, 3));
    ACHK(radix_alloc(c->rs, S));
#undef ACHK
    (void)hipMemsetAsync(d.err, 0, sizeof(uint32_t), c->stream);
    (void)hipMemsetAsync(d.steps_done, 0, sizeof(int64_t), c->stream);
    (void)hipMemsetAsync(d.scnt, 0, sizeof(int32_t) * S, c->stream);
    (void)hipMemsetAsync(d.qn, 0, sizeof(int32_t) * NQUEUE * XCD_COUNT * QN_STRIDE, c->stream);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, d.rebuild, 2, 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, d.rebuild + 2, 1, 0);
    (void)hipMemsetAsync(d.dirty, 0, (S + 3) / 4 * 4, c->stream);
    (void)hipMemsetAsync(d.wnz, 1, S, c->stream);
    (void)hipMemsetAsync(d.stats, 0, sizeof(int64_t) * 4 * max_members, c->stream);
    (void)hipMemsetAsync(d.tcur, 0, sizeof(int32_t) * max_members, c->stream);
    hipLaunchKernelGGL(k_fill_i32, dim3((S + 255) / 256), dim3(256), 0, c->stream, d.member_of, S, -1);
    for (int k = 0; k < 2; ++k) {
        (void)hipMemsetAsync(c->hcnt[k], 0, sizeof(int32_t) * S, c->stream);
        (void)hipMemsetAsync(c->bcnt[k], 0, sizeof(int32_t) * S, c->stream);
    }
    if (ensure_cells(c, 1024) != GODAR_OK) { godar_destroy(c); return GODAR_ERR_NOMEM; }
    if (upload_tables(c) != GODAR_OK || hipStreamSynchronize(c->stream) != hipSuccess) {
        godar_destroy(c);
        return GODAR_ERR_HIP;
    }
    *out = c;
    return GODAR_OK;
}

int godar_destroy(godar_ctx* c) {
    if (!c) return GODAR_ERR_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    DS& d = c->d;
    double* dbl[] = {d.wbc, d.acc, d.cur, d.fout};
    for (double* p : dbl) dfree(p);
    dfree(d.rec); dfree(d.geo); dfree(d.shel); dfree(d.kin); dfree(d.pos_of); dfree(d.pos_id);
    dfree(d.member_of); dfree(d.gid_of); dfree(d.sid_of); dfree(d.members);
    DevParams* pp = const_cast<DevParams*>(d.params); dfree(pp);
    WindGridDesc* wg = const_cast<WindGridDesc*>(d.wgrids); dfree(wg);
    double* wp = const_cast<double*>(d.wpool); dfree(wp);
    dfree(d.ckey); dfree(d.cval); dfree(d.ocnt); dfree(d.icnt);
    dfree(d.crange);
    for (int k = 0; k < 2; ++k) { dfree(c->hcnt[k]); dfree(c->bcnt[k]); }
    {
        PairLists L = lists_of(c);
        free_lists(L);
        set_lists(c, L);
    }
    dfree(d.wout); dfree(d.wflag); dfree(d.cls); dfree(d.qlist); dfree(d.qn); dfree(d.bgat); dfree(d.seeds); dfree(d.nseeds); dfree(d.heap); dfree(d.etab);
    dfree(d.netab); dfree(d.wtab); dfree(d.nwtab); dfree(d.ovw); dfree(d.dirty); dfree(d.comp);
    dfree(d.wnz); dfree(d.stats); dfree(d.tcur); dfree(d.err); dfree(d.steps_done);
    dfree(d.sup); dfree(d.scnt); dfree(d.sout); dfree(d.soff); dfree(d.sblk); dfree(d.x0); dfree(d.rebuild);
    radix_free(c->rs);
    collect_events(c);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->ejoin) (void)hipEventDestroy(c->ejoin);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->efork) (void)hipEventDestroy(c->efork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GODAR_OK;
}

const char* godar_last_error(const godar_ctx* c) { return c ? c->err.c_str() : "null context"; }

// Device bytes per floe slot: every array godar_create and alloc_lists size
// by the floe capacity (radix scratch included), at the context's current
// list capacities.
static int64_t bytes_per_floe(const godar_ctx* c) {
    int64_t per = 0;
    per += sizeof(FloeRec) + sizeof(GeoRec) + sizeof(ShelRec) + sizeof(KinRec);
    per += 4 + 4 /*pos_of, pos_id*/ + 4 * 8 /*wbc*/ + 4 /*member_of*/ + 4 + 4 /*ckey, cval*/;
    per += 4 + 4 /*ocnt, icnt*/ + 2 * (4 + 4) /*hcnt, bcnt*/;
    per += (int64_t)c->CO * (8 + 8 + 4 + 4 + 8 + 8) + (int64_t)c->CI * (8 + 4);   // oj od evf ov ssc ssb; ik ibond
    per += 2 * ((int64_t)c->HC * sizeof(HistRec) + (int64_t)c->BC * sizeof(BondRec));
    per += (int64_t)c->BC * sizeof(BondTerms) + 1 /*bgat*/;
    per += NACC * 8 + sizeof(WallOut) + 1 /*wflag*/ + 1 /*cls*/ + NQUEUE * 4 /*qlist*/;
    per += (int64_t)c->SEEDK * 8 + (int64_t)c->HEAPK * 8 + (int64_t)c->TABK * sizeof(EventOut);
    per += sizeof(WallOut) /*wtab*/ + 4 /*ovw*/ + 7 * 8 /*cur*/ + 1 + 1 /*dirty, wnz*/ + 4 /*comp*/;
    per += 33 * 8 /*fout*/ + 4 + 4 /*radix keys, values*/;
    per += (c->S > 0 ? c->sup_cap / c->S : 0) * 8 /*superset pool share*/ + 4 + 4 + 8 + 16 /*scnt sout soff x0*/;
    per += 1 /*sblk, per SCAN_TILE floes*/;
    return per;
}

// Search shape of a member of n floes whose largest radius is rmax (radii
// only shrink, by ridging): the superset skin is skin_frac (rmax + rtree);
// the member is dense when the 3x3 cells around a floe on a grid of cells
// >= rmax + rtree + skin hold dense_scan floes or more (on average); the
// cells are >= (rmax + rtree + skin) / span, span = 2 for the dense search
// and thread_span for the thread search, so that the (2 span + 1)^2 cells
// around a floe hold its whole superset.  Grids are coarsened to at most
// 4n + 64 cells.
void member_search(const godar_ctx* c, DevParams& dp, int n, double rmax, int64_t& gx, int64_t& gy) {
    dp.skin = c->skin_frac * (rmax + dp.rtree);
    auto grid = [&](int span) {
        dp.cell = (rmax + dp.rtree + dp.skin) * 1.0000001 / span + 1e-9;
        gx = std::max<int64_t>(1, (int64_t)std::floor(dp.nx / dp.cell));
        gy = std::max<int64_t>(1, (int64_t)std::floor(dp.ny / dp.cell));
        while (gx * gy > 4 * (int64_t)n + 64) { gx = (gx + 1) / 2; gy = (gy + 1) / 2; dp.cell *= 2; }
    };
    grid(1);
    dp.dense = 9.0 * (double)n / (double)(gx * gy) >= (double)c->dense_scan ? 1 : 0;
    dp.span = dp.dense ? 2 : c->thread_span;
    if (dp.span != 1) grid(dp.span);
    dp.gx = (int32_t)gx;
    dp.gy = (int32_t)gy;
}

// Every member's search shape anew from the context's settings (skin,
// thread span, dense threshold), with cell ranges laid out one after the
// other; the next step builds the supersets.  Nothing changes if the cell
// table cannot grow.
int research_all(godar_ctx* c) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<DevParams> dps = c->dparams;
    std::vector<int32_t> cn(c->Mcap, 0);
    int64_t base = 0;
    for (int m = 0; m < c->Mcap; ++m) {
        if (!c->members[m].active) continue;
        int64_t gx, gy;
        member_search(c, dps[m], c->members[m].n, c->rmax[m], gx, gy);
        dps[m].cell_base = (int32_t)base;
        cn[m] = (int32_t)(gx * gy);
        base += gx * gy;
        if (base >= ((int64_t)1 << 30)) return fail(c, GODAR_ERR_NOMEM, "search: too many cells");
    }
    int rc = ensure_cells(c, (int32_t)base);
    if (rc != GODAR_OK) return rc;
    c->dparams = dps;
    c->cell_n = cn;
    rc = upload_tables(c);
    if (rc != GODAR_OK) return rc;
    force_rebuild(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GODAR_OK;
}

int godar_member_footprint(const godar_ctx* c, int32_t n, int64_t* bytes) {
    if (!c || !bytes || n < 0) return GODAR_ERR_ARG;
    *bytes = bytes_per_floe(c) * n;
    return GODAR_OK;
}

int godar_capacity_left(const godar_ctx* c, int64_t* floes, int32_t* members) {
    if (!c) return GODAR_ERR_ARG;
    int64_t used = 0;
    int32_t nm = 0;
    for (const Member& m : c->members)
        if (m.active) { used += m.n; ++nm; }
    if (floes) *floes = c->S - used;
    if (members) *members = c->Mcap - nm;
    return GODAR_OK;
}

int godar_member_add(godar_ctx* c, const godar_params* p, int32_t n, const double* x,
                     const double* y, const double* u, const double* v, const double* theta,
                     const double* omega, const double* r, const double* h, int32_t tstep0,
                     int32_t* member) {
    if (!c || !p || n <= 0 || !x || !y || !u || !v || !theta || !omega || !r || !h || !member)
        return fail(c, GODAR_ERR_ARG, "member_add: bad argument");
    if (n > 0xFFFF * 64) return fail(c, GODAR_ERR_ARG, "member_add: member too large");
    if (!(p->dt > 0) || !(p->rtree >= 0) || !(p->nx > 0) || !(p->ny > 0))
        return fail(c, GODAR_ERR_ARG, "member_add: dt, nx, ny must be positive, rtree >= 0");
    // The reference searches a tree of the positions at the last rebuild,
    // rebuilt when tstep == 1 or mod(tstep, ntree) == 0 (stepper.f90:44-50);
    // the search here always uses the current positions, which is ntree = 1
    // (the default, par_get.f90:28).
    if (p->ntree != 1)
        return fail(c, GODAR_ERR_ARG, "member_add: ntree must be 1 (the neighbour search uses the "
                                      "current positions every step)");
    int m = -1;
    for (int k = 0; k < c->Mcap; ++k)
        if (!c->members[k].active) { m = k; break; }
    if (m < 0) return fail(c, GODAR_ERR_NOMEM, "member_add: no free member slot");
    std::vector<Range> used;
    for (const Member& mm : c->members)
        if (mm.active) used.push_back({mm.first, mm.n});
    int64_t first;
    if (!find_range(c, n, c->S, used, first))
        return fail(c, GODAR_ERR_NOMEM, "member_add: no room for the floes");
    // cell grid: cell >= max query radius (r only shrinks, by ridging)
    double rmax = 0;
    for (int i = 0; i < n; ++i) {
        if (!(r[i] > 0) || !(h[i] > 0) || !std::isfinite(x[i]) || !std::isfinite(y[i]))
            return fail(c, GODAR_ERR_ARG, "member_add: radii/thicknesses must be positive and positions finite");
        rmax = std::max(rmax, r[i]);
    }
    DevParams dp;
    derive_params(*p, dp);
    int64_t gx, gy;
    member_search(c, dp, n, rmax, gx, gy);
    int32_t ncell = (int32_t)(gx * gy);
    std::vector<Range> cused;
    for (int k = 0; k < c->Mcap; ++k)
        if (c->members[k].active) cused.push_back({c->dparams[k].cell_base, c->cell_n[k]});
    int64_t cbase;
    find_range(c, ncell, (int64_t)1 << 30, cused, cbase);
    int rc = ensure_cells(c, (int32_t)(cbase + ncell));
    if (rc != GODAR_OK) return rc;
    dp.cell_base = (int32_t)cbase;
    c->cell_n[m] = ncell;
    // storage order: the member's floes sorted (stably) by their first cell,
    // as the step's cell-ordered sweep will visit them
    std::vector<int32_t> order(n);
    {
        std::vector<uint32_t> key(n);
        auto coord = [](double v, double cell, int g) {
            double q = std::floor(v / cell);
            if (!(q >= 0.0)) return 0;
            if (q >= (double)(g - 1)) return g - 1;
            return (int)q;
        };
        for (int i = 0; i < n; ++i)
            key[i] = (uint32_t)(coord(y[i], dp.cell, dp.gy) * dp.gx + coord(x[i], dp.cell, dp.gx));
        for (int i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    }
    // floe state and floe_properties (physics.f90:179-197), as ini_get does
    std::vector<FloeRec> recs(n);
    std::vector<int32_t> gid(n), sid(n);
    for (int t = 0; t < n; ++t) {
        const int i = order[t];
        FloeRec& q = recs[t];
        q.x = x[i]; q.y = y[i]; q.u = u[i]; q.v = v[i]; q.omega = omega[i]; q.theta = theta[i];
        q.r = r[i]; q.h = h[i];
        q.mass = dp.rhoice * dp.pi * h[i] * (r[i] * r[i]);
        q.inertia = 0.5 * q.mass * (r[i] * r[i]);
        q.hfa = h[i] * (dp.rhowater - dp.rhoice) / dp.rhowater;
        q.hfw = h[i] * dp.rhoice / dp.rhowater;
        gid[t] = (int32_t)(first + i);
        sid[i] = (int32_t)(first + t);
    }
    spill(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    DS& d = c->d;
    HIPCHK(c, hipMemcpy(d.rec + first, recs.data(), sizeof(FloeRec) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.gid_of + first, gid.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.sid_of + first, sid.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    c->perm[m] = order;
    for (int k = 0; k < 4; ++k) HIPCHK(c, hipMemset(d.wbc + k * c->S + first, 0, sizeof(double) * n));
    for (int k = 0; k < 2; ++k) {
        HIPCHK(c, hipMemset(c->hcnt[k] + first, 0, sizeof(int32_t) * n));
        HIPCHK(c, hipMemset(c->bcnt[k] + first, 0, sizeof(int32_t) * n));
    }
    HIPCHK(c, hipMemset(d.dirty + first, 0, n));
    HIPCHK(c, hipMemset(d.wnz + first, 0, n));        // wall history rows zeroed above
    std::vector<int32_t> mo(n, m);
    HIPCHK(c, hipMemcpy(d.member_of + first, mo.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    c->dparams[m] = dp;
    c->rmax[m] = rmax;
    c->members[m].first = (int32_t)first;
    c->members[m].n = n;
    c->members[m].tstep = tstep0;
    c->members[m].active = 1;
    c->members[m].tstop = 0;
    c->N_used = std::max<int32_t>(c->N_used, (int32_t)(first + n));
    rc = upload_tables(c);
    if (rc != GODAR_OK) return rc;
    force_rebuild(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *member = m;
    return GODAR_OK;
}

// ---------------------------------------------------------------- checkpoint
// Block layout (native byte order), everything in the caller's floe order:
//   CkHeader | godar_params | n x 12 doubles floe state (x y u v omega r h
//   mass inertia hfa hfw theta) | n x 4 doubles wall histories |
//   n int32 history counts | histories (int32 partner, int32 0, 4 doubles) |
//   n int32 bond counts | bonds (int32 partner, int32 bond, 10 doubles)
// Partners are member-local indices.
namespace {
struct CkHeader {
    char magic[8];
    int32_t version, n, tstep, precision;
    int64_t nhist, nbond;
};
constexpr int32_t CK_VERSION = 1;
constexpr size_t CK_HIST = 8 + 4 * 8, CK_BOND = 8 + 10 * 8;

struct MemberState {
    std::vector<FloeRec> rec;        // caller's order
    std::vector<double> wbc;         // 4 per floe
    std::vector<int32_t> hcnt, bcnt;
    std::vector<HistRec> hist;       // j as member-local index
    std::vector<BondRec> bond;
};

int read_member_state(godar_ctx* c, int32_t m, MemberState& out) {
    spill(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const Member& mm = c->members[m];
    const int n = mm.n;
    const int64_t S = c->S;
    const int o = c->parity;
    const std::vector<int32_t>& pm = c->perm[m];
    std::vector<FloeRec> rs(n);
    HIPCHK(c, hipMemcpy(rs.data(), c->d.rec + mm.first, sizeof(FloeRec) * n, hipMemcpyDeviceToHost));
    out.rec.assign(n, FloeRec{});
    for (int t = 0; t < n; ++t) out.rec[pm[t]] = rs[t];
    out.wbc.assign((size_t)4 * n, 0.0);
    std::vector<double> row(n);
    for (int k = 0; k < 4; ++k) {
        HIPCHK(c, hipMemcpy(row.data(), c->d.wbc + k * S + mm.first, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (int t = 0; t < n; ++t) out.wbc[(size_t)4 * pm[t] + k] = row[t];
    }
    std::vector<int32_t> hc(n), bc(n);
    HIPCHK(c, hipMemcpy(hc.data(), c->hcnt[o] + mm.first, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(bc.data(), c->bcnt[o] + mm.first, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    std::vector<std::vector<HistRec>> hl(n);
    std::vector<std::vector<BondRec>> bl(n);
    std::vector<HistRec> hrow(n);
    for (int q = 0; q < c->HC; ++q) {
        bool any = false;
        for (int t = 0; t < n; ++t) any |= q < hc[t];
        if (!any) break;
        HIPCHK(c, hipMemcpy(hrow.data(), c->hist[o] + q * S + mm.first, sizeof(HistRec) * n, hipMemcpyDeviceToHost));
        for (int t = 0; t < n; ++t)
            if (q < hc[t] && hrow[t].j >= 0) {
                HistRec h = hrow[t];
                h.j -= mm.first;
                hl[pm[t]].push_back(h);
            }
    }
    std::vector<BondRec> brow(n);
    for (int q = 0; q < c->BC; ++q) {
        bool any = false;
        for (int t = 0; t < n; ++t) any |= q < bc[t];
        if (!any) break;
        HIPCHK(c, hipMemcpy(brow.data(), c->bond[o] + q * S + mm.first, sizeof(BondRec) * n, hipMemcpyDeviceToHost));
        for (int t = 0; t < n; ++t)
            if (q < bc[t] && brow[t].j >= 0) {
                BondRec b = brow[t];
                b.j -= mm.first;
                bl[pm[t]].push_back(b);
            }
    }
    out.hcnt.assign(n, 0);
    out.bcnt.assign(n, 0);
    out.hist.clear();
    out.bond.clear();
    for (int i = 0; i < n; ++i) {
        out.hcnt[i] = (int32_t)hl[i].size();
        out.bcnt[i] = (int32_t)bl[i].size();
        out.hist.insert(out.hist.end(), hl[i].begin(), hl[i].end());
        out.bond.insert(out.bond.end(), bl[i].begin(), bl[i].end());
    }
    return GODAR_OK;
}
}  // namespace

int godar_checkpoint_size(godar_ctx* c, int32_t m, int64_t* bytes) {
    if (!valid_member(c, m) || !bytes) return fail(c, GODAR_ERR_ARG, "checkpoint_size: bad argument");
    MemberState st;
    int rc = read_member_state(c, m, st);
    if (rc != GODAR_OK) return rc;
    const int64_t n = c->members[m].n;
    *bytes = (int64_t)sizeof(CkHeader) + (int64_t)sizeof(godar_params) + n * 12 * 8 + n * 4 * 8 +
             n * 4 + (int64_t)st.hist.size() * (int64_t)CK_HIST + n * 4 +
             (int64_t)st.bond.size() * (int64_t)CK_BOND;
    return GODAR_OK;
}

int godar_checkpoint_save(godar_ctx* c, int32_t m, const godar_params* p, void* buf, int64_t cap) {
    if (!valid_member(c, m) || !p || !buf) return fail(c, GODAR_ERR_ARG, "checkpoint_save: bad argument");
    int rc = download_members(c);
    if (rc != GODAR_OK) return rc;
    MemberState st;
    rc = read_member_state(c, m, st);
    if (rc != GODAR_OK) return rc;
    const int n = c->members[m].n;
    int64_t need = 0;
    rc = godar_checkpoint_size(c, m, &need);
    if (rc != GODAR_OK) return rc;
    if (cap < need) return fail(c, GODAR_ERR_ARG, "checkpoint_save: buffer too small");
    char* w = static_cast<char*>(buf);
    CkHeader h{};
    std::memcpy(h.magic, "GODARCK1", 8);
    h.version = CK_VERSION; h.n = n; h.tstep = c->members[m].tstep; h.precision = c->precision;
    h.nhist = (int64_t)st.hist.size(); h.nbond = (int64_t)st.bond.size();
    std::memcpy(w, &h, sizeof(h)); w += sizeof(h);
    std::memcpy(w, p, sizeof(godar_params)); w += sizeof(godar_params);
    for (int i = 0; i < n; ++i) {
        const FloeRec& q = st.rec[i];
        double v[12] = {q.x, q.y, q.u, q.v, q.omega, q.r, q.h, q.mass, q.inertia, q.hfa, q.hfw, q.theta};
        std::memcpy(w, v, sizeof(v)); w += sizeof(v);
    }
    std::memcpy(w, st.wbc.data(), sizeof(double) * 4 * n); w += sizeof(double) * 4 * n;
    std::memcpy(w, st.hcnt.data(), sizeof(int32_t) * n); w += sizeof(int32_t) * n;
    for (const HistRec& r : st.hist) {
        int32_t ids[2] = {r.j, 0};
        double v[4] = {r.deltat, r.thetarelc, r.fcn, r.fct};
        std::memcpy(w, ids, 8); w += 8;
        std::memcpy(w, v, sizeof(v)); w += sizeof(v);
    }
    std::memcpy(w, st.bcnt.data(), sizeof(int32_t) * n); w += sizeof(int32_t) * n;
    for (const BondRec& r : st.bond) {
        int32_t ids[2] = {r.j, r.bond};
        double v[10] = {r.damage, r.deltanb, r.deltatb, r.trb_ji, r.trb_ij, r.rb, r.hb, r.lb, r.sb, r.ib};
        std::memcpy(w, ids, 8); w += 8;
        std::memcpy(w, v, sizeof(v)); w += sizeof(v);
    }
    return GODAR_OK;
}

int godar_member_restore(godar_ctx* c, const void* buf, int64_t size, int32_t* member) {
    if (!c || !buf || !member || size < (int64_t)(sizeof(CkHeader) + sizeof(godar_params)))
        return fail(c, GODAR_ERR_ARG, "member_restore: bad argument");
    const char* r = static_cast<const char*>(buf);
    const char* end = r + size;
    CkHeader h;
    std::memcpy(&h, r, sizeof(h)); r += sizeof(h);
    if (std::memcmp(h.magic, "GODARCK1", 8) != 0 || h.version != CK_VERSION || h.n <= 0 ||
        h.nhist < 0 || h.nbond < 0)
        return fail(c, GODAR_ERR_ARG, "member_restore: not a checkpoint block");
    if (h.precision != c->precision)
        return fail(c, GODAR_ERR_ARG, "member_restore: the block was saved in a context of the other "
                                      "precision (it would not continue exactly)");
    const int n = h.n;
    const int64_t need = (int64_t)sizeof(CkHeader) + (int64_t)sizeof(godar_params) + (int64_t)n * 12 * 8 +
                         (int64_t)n * 4 * 8 + (int64_t)n * 4 + h.nhist * (int64_t)CK_HIST +
                         (int64_t)n * 4 + h.nbond * (int64_t)CK_BOND;
    if (size < need) return fail(c, GODAR_ERR_ARG, "member_restore: block truncated");
    godar_params p;
    std::memcpy(&p, r, sizeof(p)); r += sizeof(p);
    std::vector<double> x(n), y(n), u(n), v(n), om(n), rr(n), hh(n), th(n);
    for (int i = 0; i < n; ++i) {
        double q[12];
        std::memcpy(q, r, sizeof(q)); r += sizeof(q);
        x[i] = q[0]; y[i] = q[1]; u[i] = q[2]; v[i] = q[3]; om[i] = q[4]; rr[i] = q[5]; hh[i] = q[6];
        th[i] = q[11];
    }
    std::vector<double> wbc((size_t)4 * n);
    std::memcpy(wbc.data(), r, sizeof(double) * 4 * n); r += sizeof(double) * 4 * n;
    std::vector<int32_t> hc(n), bc(n);
    std::memcpy(hc.data(), r, sizeof(int32_t) * n); r += sizeof(int32_t) * n;
    const char* hist_at = r;
    r += h.nhist * (int64_t)CK_HIST;
    std::memcpy(bc.data(), r, sizeof(int32_t) * n); r += sizeof(int32_t) * n;
    const char* bond_at = r;
    (void)end;
    int maxh = 0, maxb = 0;
    int64_t sh = 0, sb = 0;
    for (int i = 0; i < n; ++i) {
        if (hc[i] < 0 || bc[i] < 0) return fail(c, GODAR_ERR_ARG, "member_restore: bad counts");
        maxh = std::max(maxh, hc[i]); maxb = std::max(maxb, bc[i]);
        sh += hc[i]; sb += bc[i];
    }
    if (sh != h.nhist || sb != h.nbond) return fail(c, GODAR_ERR_ARG, "member_restore: bad counts");
    // every record is owned by the lower index of its pair: partner in (i, n)
    {
        const char* q = hist_at;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < hc[i]; ++k, q += CK_HIST) {
                int32_t j;
                std::memcpy(&j, q, 4);
                if (j <= i || j >= n) return fail(c, GODAR_ERR_ARG, "member_restore: history partner out of range");
            }
        q = bond_at;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < bc[i]; ++k, q += CK_BOND) {
                int32_t j;
                std::memcpy(&j, q, 4);
                if (j <= i || j >= n) return fail(c, GODAR_ERR_ARG, "member_restore: bond partner out of range");
            }
    }
    if (maxh > c->HC || maxb > c->BC) {
        int rc = godar_reserve(c, 0, 0, std::max(maxh, c->HC), std::max(maxb, c->BC));
        if (rc != GODAR_OK) return rc;
    }
    int32_t m = -1;
    int rc = godar_member_add(c, &p, n, x.data(), y.data(), u.data(), v.data(), th.data(), om.data(),
                              rr.data(), hh.data(), h.tstep, &m);
    if (rc != GODAR_OK) return rc;
    // the floe properties that ridging changed are the recomputed ones
    // (floe_properties of r and h); the pair and wall state goes to storage
    const Member& mm = c->members[m];
    const std::vector<int32_t>& pm = c->perm[m];
    const int64_t S = c->S;
    const int o = c->parity;
    std::vector<int32_t> slot(n);
    for (int t = 0; t < n; ++t) slot[pm[t]] = t;
    std::vector<double> row(n);
    for (int k = 0; k < 4; ++k) {
        for (int i = 0; i < n; ++i) row[slot[i]] = wbc[(size_t)4 * i + k];
        HIPCHK(c, hipMemcpy(c->d.wbc + k * S + mm.first, row.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMemset(c->d.wnz + mm.first, 1, n));   // restored rows may be nonzero
    std::vector<int32_t> cnt(n);
    for (int i = 0; i < n; ++i) cnt[slot[i]] = hc[i];
    HIPCHK(c, hipMemcpy(c->hcnt[o] + mm.first, cnt.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    for (int i = 0; i < n; ++i) cnt[slot[i]] = bc[i];
    HIPCHK(c, hipMemcpy(c->bcnt[o] + mm.first, cnt.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    {
        std::vector<HistRec> rows((size_t)std::max(maxh, 1) * n);
        const char* q = hist_at;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < hc[i]; ++k) {
                int32_t ids[2];
                double vv[4];
                std::memcpy(ids, q, 8); q += 8;
                std::memcpy(vv, q, sizeof(vv)); q += sizeof(vv);
                HistRec& hr = rows[(size_t)k * n + slot[i]];
                hr.j = mm.first + ids[0]; hr.pad_ = 0;
                hr.deltat = vv[0]; hr.thetarelc = vv[1]; hr.fcn = vv[2]; hr.fct = vv[3];
            }
        for (int k = 0; k < maxh; ++k)
            HIPCHK(c, hipMemcpy(c->hist[o] + k * S + mm.first, rows.data() + (size_t)k * n,
                                sizeof(HistRec) * n, hipMemcpyHostToDevice));
    }
    {
        std::vector<BondRec> rows((size_t)std::max(maxb, 1) * n);
        const char* q = bond_at;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < bc[i]; ++k) {
                int32_t ids[2];
                double vv[10];
                std::memcpy(ids, q, 8); q += 8;
                std::memcpy(vv, q, sizeof(vv)); q += sizeof(vv);
                BondRec& b = rows[(size_t)k * n + slot[i]];
                b.j = mm.first + ids[0]; b.bond = ids[1];
                b.damage = vv[0]; b.deltanb = vv[1]; b.deltatb = vv[2]; b.trb_ji = vv[3]; b.trb_ij = vv[4];
                b.rb = vv[5]; b.hb = vv[6]; b.lb = vv[7]; b.sb = vv[8]; b.ib = vv[9];
            }
        for (int k = 0; k < maxb; ++k)
            HIPCHK(c, hipMemcpy(c->bond[o] + k * S + mm.first, rows.data() + (size_t)k * n,
                                sizeof(BondRec) * n, hipMemcpyHostToDevice));
    }
    *member = m;
    return GODAR_OK;
}

int godar_member_remove(godar_ctx* c, int32_t m) {
    if (!valid_member(c, m)) return fail(c, GODAR_ERR_ARG, "member_remove: bad member");
    spill(c);                       // before the floe count changes
    int rc = download_members(c);   // keep device-side step counters
    if (rc != GODAR_OK) return rc;
    Member& mm = c->members[m];
    hipLaunchKernelGGL(k_fill_i32, dim3((mm.n + 255) / 256), dim3(256), 0, c->stream,
                       c->d.member_of + mm.first, (int64_t)mm.n, -1);
    mm.active = 0;
    c->cell_n[m] = 0;
    int32_t hw = 0;
    for (const Member& q : c->members)
        if (q.active) hw = std::max(hw, q.first + q.n);
    c->N_used = hw;
    rc = upload_members_only(c);
    if (rc != GODAR_OK) return rc;
    force_rebuild(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GODAR_OK;
}

int godar_set_forcing(godar_ctx* c, int32_t m, double ua, double va, double uw, double vw) {
    if (!valid_member(c, m)) return fail(c, GODAR_ERR_ARG, "set_forcing: bad member");
    DevParams& d = c->dparams[m];
    d.ua = ua; d.va = va; d.uw = uw; d.vw = vw;
    HIPCHK(c, hipMemcpyAsync((void*)(c->d.params + m), &d, sizeof(DevParams), hipMemcpyHostToDevice,
                             c->stream));
    force_refresh(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GODAR_OK;
}

int godar_set_wind_frames(godar_ctx* c, int32_t m, int32_t nxn, int32_t nyn, double x0, double y0,
                          double dx, double dy, int32_t nframes, double t0, double frame_dt,
                          int32_t disc_average, const double* ua, const double* va) {
    if (!valid_member(c, m) || nxn < 2 || nyn < 2 || !(dx > 0) || !(dy > 0) || !ua || !va ||
        nframes < 1 || !std::isfinite(t0) || (nframes > 1 && !(frame_dt > 0)))
        return fail(c, GODAR_ERR_ARG, "set_wind_frames: bad argument");
    WindGridDesc g;
    g.nxn = nxn; g.nyn = nyn; g.x0 = x0; g.y0 = y0; g.dx = dx; g.dy = dy;
    g.offset = (int64_t)c->wpool.size();
    g.nframes = nframes; g.avg = disc_average ? 1 : 0;
    g.t0 = t0; g.fdt = nframes > 1 ? frame_dt : 1.0;
    const int64_t cnt = (int64_t)nxn * nyn;
    for (int32_t f = 0; f < nframes; ++f) {
        c->wpool.insert(c->wpool.end(), ua + f * cnt, ua + (f + 1) * cnt);
        c->wpool.insert(c->wpool.end(), va + f * cnt, va + (f + 1) * cnt);
    }
    c->wgrids.push_back(g);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((int64_t)c->wpool.size() > c->wpool_dev_cap) {
        double* wp = nullptr;
        const int64_t cap = std::max<int64_t>((int64_t)c->wpool.size(), 2 * c->wpool_dev_cap);
        if (dalloc(&wp, cap) != hipSuccess) {
            (void)hipGetLastError();
            c->wgrids.pop_back();
            c->wpool.resize((size_t)g.offset);
            return fail(c, GODAR_ERR_NOMEM, "set_wind_frames: out of device memory");
        }
        double* old = const_cast<double*>(c->d.wpool);
        dfree(old);
        c->d.wpool = wp;
        c->wpool_dev_cap = cap;
    }
    if ((int)c->wgrids.size() > c->wgrids_dev_cap) {
        WindGridDesc* wg = nullptr;
        const int cap = std::max<int>((int)c->wgrids.size(), 2 * c->wgrids_dev_cap);
        if (dalloc(&wg, cap) != hipSuccess) {
            (void)hipGetLastError();
            c->wgrids.pop_back();
            c->wpool.resize((size_t)g.offset);
            return fail(c, GODAR_ERR_NOMEM, "set_wind_frames: out of device memory");
        }
        WindGridDesc* old = const_cast<WindGridDesc*>(c->d.wgrids);
        dfree(old);
        c->d.wgrids = wg;
        c->wgrids_dev_cap = cap;
    }
    HIPCHK(c, hipMemcpy((void*)c->d.wpool, c->wpool.data(), sizeof(double) * c->wpool.size(),
                        hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy((void*)c->d.wgrids, c->wgrids.data(), sizeof(WindGridDesc) * c->wgrids.size(),
                        hipMemcpyHostToDevice));
    c->dparams[m].wind_grid = (int32_t)c->wgrids.size() - 1;
    HIPCHK(c, hipMemcpy((void*)(c->d.params + m), &c->dparams[m], sizeof(DevParams), hipMemcpyHostToDevice));
    force_refresh(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GODAR_OK;
}

int godar_set_wind_grid(godar_ctx* c, int32_t m, int32_t nxn, int32_t nyn, double x0, double y0,
                        double dx, double dy, const double* ua, const double* va) {
    return godar_set_wind_frames(c, m, nxn, nyn, x0, y0, dx, dy, 1, 0.0, 1.0, 0, ua, va);
}

int godar_set_bonds(godar_ctx* c, int32_t m, int32_t nb, const int32_t* bi, const int32_t* bj,
                    const double* st) {
    if (!valid_member(c, m) || nb < 0 || (nb > 0 && (!bi || !bj || !st)))
        return fail(c, GODAR_ERR_ARG, "set_bonds: bad argument");
    const Member& mm = c->members[m];
    int64_t S = c->S;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int o = c->parity;
    std::vector<int32_t> cnt(mm.n, 0), slot_of(mm.n);
    for (int t = 0; t < mm.n; ++t) slot_of[c->perm[m][t]] = t;   // caller's index -> storage offset
    std::vector<BondRec> recs((int64_t)mm.n * c->BC);
    for (int k = 0; k < nb; ++k) {
        int i = bi[k], j = bj[k];
        if (i < 0 || j <= i || j >= mm.n) return fail(c, GODAR_ERR_ARG, "set_bonds: need 0 <= i < j < n");
        const int t = slot_of[i];
        if (cnt[t] >= c->BC) return fail(c, GODAR_ERR_CAPACITY, "set_bonds: bond list full, raise it with godar_reserve");
        BondRec& b = recs[(int64_t)cnt[t] * mm.n + t];
        const double* s = st + 11 * k;
        b.j = mm.first + j; b.bond = (int32_t)s[0]; b.damage = s[1]; b.deltanb = s[2]; b.deltatb = s[3];
        b.trb_ji = s[4]; b.trb_ij = s[5]; b.rb = s[6]; b.hb = s[7]; b.lb = s[8]; b.sb = s[9]; b.ib = s[10];
        cnt[t]++;
    }
    for (int q = 0; q < c->BC; ++q)
        HIPCHK(c, hipMemcpy(c->bond[o] + q * S + mm.first, recs.data() + (int64_t)q * mm.n,
                            sizeof(BondRec) * mm.n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->bcnt[o] + mm.first, cnt.data(), sizeof(int32_t) * mm.n, hipMemcpyHostToDevice));
    return GODAR_OK;
}

int godar_reserve(godar_ctx* c, int32_t outer, int32_t inner, int32_t histc, int32_t bonds) {
    if (!c || outer < 0 || inner < 0 || histc < 0 || bonds < 0 || outer > 0xFFFE ||
        inner > (1 << 15) || histc > (1 << 12) || bonds > (1 << 12))
        return fail(c, GODAR_ERR_ARG, "reserve: bad argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int nCO = outer ? outer : c->CO, nCI = inner ? inner : c->CI;
    const int nHC = histc ? histc : c->HC, nBC = bonds ? bonds : c->BC;
    if (nCO == c->CO && nCI == c->CI && nHC == c->HC && nBC == c->BC) return GODAR_OK;
    const int64_t S = c->S;
    const int o = c->parity;
    // The kernels walk a floe's records up to its count, and the last step's
    // candidate lists stay readable (godar_get_pairs): a list may shrink only
    // down to the longest one in use.
    auto longest = [&](const int32_t* cnt, int32_t& mx) -> int {
        mx = 0;
        if (c->N_used == 0) return GODAR_OK;
        std::vector<int32_t> h(c->N_used);
        HIPCHK(c, hipMemcpy(h.data(), cnt, sizeof(int32_t) * c->N_used, hipMemcpyDeviceToHost));
        for (int32_t v : h) mx = std::max(mx, v);
        return GODAR_OK;
    };
    int32_t mo = 0, mi = 0, mh = 0, mb = 0;
    int rc;
    if (nCO < c->CO && (rc = longest(c->d.ocnt, mo)) != GODAR_OK) return rc;
    if (nCI < c->CI && (rc = longest(c->d.icnt, mi)) != GODAR_OK) return rc;
    if (nHC < c->HC && (rc = longest(c->hcnt[o], mh)) != GODAR_OK) return rc;
    if (nBC < c->BC && (rc = longest(c->bcnt[o], mb)) != GODAR_OK) return rc;
    if (nCO < mo || nCI < mi || nHC < mh || nBC < mb)
        return fail(c, GODAR_ERR_ARG,
                    "reserve: a capacity is below the entries a floe holds (outer " + std::to_string(mo) +
                        ", inner " + std::to_string(mi) + ", histories " + std::to_string(mh) +
                        ", bonds " + std::to_string(mb) + ")");
    // New lists first: if an allocation fails the context keeps its lists
    // and stays usable.
    PairLists nl;
    if (alloc_lists(nl, S, nCO, nCI, nHC, nBC) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, GODAR_ERR_NOMEM, "reserve: out of device memory (the current capacities stay)");
    }
    PairLists ol = lists_of(c);
    // slot-major lists: the first rows of each carry over
    auto rows = [&](void* dst, const void* src, size_t elem, int oc, int nc) {
        return hipMemcpy(dst, src, elem * (size_t)S * (size_t)std::min(oc, nc), hipMemcpyDeviceToDevice);
    };
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t x) { e = x; return x == hipSuccess; };
    bool good = ok(rows(nl.oj, ol.oj, sizeof(int64_t), c->CO, nCO)) &&
                ok(rows(nl.od, ol.od, sizeof(double), c->CO, nCO)) &&
                ok(rows(nl.evf, ol.evf, sizeof(uint32_t), c->CO, nCO)) &&
                ok(rows(nl.ov, ol.ov, sizeof(int32_t), c->CO, nCO)) &&
                ok(rows(nl.ssc, ol.ssc, sizeof(double), c->CO, nCO)) &&
                ok(rows(nl.ssb, ol.ssb, sizeof(double), c->CO, nCO)) &&
                ok(rows(nl.ik, ol.ik, sizeof(int64_t), c->CI, nCI)) &&
                ok(rows(nl.ibond, ol.ibond, sizeof(int32_t), c->CI, nCI)) &&
                ok(rows(nl.bterm, ol.bterm, sizeof(BondTerms), c->BC, nBC)) &&
                ok(rows(nl.hist[o], ol.hist[o], sizeof(HistRec), c->HC, nHC)) &&
                ok(rows(nl.bond[o], ol.bond[o], sizeof(BondRec), c->BC, nBC));
    if (!good) {
        free_lists(nl);
        return fail(c, GODAR_ERR_HIP, std::string("reserve: ") + hipGetErrorString(e));
    }
    free_lists(ol);
    set_lists(c, nl);
    c->CO = nCO; c->CI = nCI; c->HC = nHC; c->BC = nBC;
    sync_ds(c);
    return GODAR_OK;
}

// stages whose per-floe kernels take a tunable block size
// (stage 6 has no kernel: the bond gather is part of the integration)
static bool tunable_stage(int st) { return st == 0 || (st >= 2 && st <= 5) || st == 7 || st >= 9; }

int godar_set_block(godar_ctx* c, int32_t threads) {
    if (!c || threads < 64 || threads > FLOE_BLOCK || (threads & 63)) return GODAR_ERR_ARG;
    c->block = threads;
    for (int k = 0; k < GODAR_NSTAGES; ++k)
        if (tunable_stage(k)) c->sblock[k] = threads;
    return GODAR_OK;
}

int godar_set_dense_search(godar_ctx* c, int32_t floes) {
    if (!c || floes < 0) return GODAR_ERR_ARG;
    c->dense_scan = floes;
    return GODAR_OK;
}

int godar_set_skin(godar_ctx* c, double frac) {
    if (!c || !(frac >= 0.0) || frac > 1.0) return GODAR_ERR_ARG;
    c->skin_frac = frac;
    return GODAR_OK;
}

int godar_set_search(godar_ctx* c, double skin_frac, int32_t thread_span, int32_t dense_scan) {
    if (!c || !(skin_frac >= 0.0) || skin_frac > 1.0 || thread_span < 1 || thread_span > 2 ||
        dense_scan < 0)
        return c ? fail(c, GODAR_ERR_ARG, "set_search: skin in [0, 1], span 1 or 2, threshold >= 0")
                 : GODAR_ERR_ARG;
    const double s0 = c->skin_frac;
    const int t0 = c->thread_span, d0 = c->dense_scan;
    c->skin_frac = skin_frac;
    c->thread_span = thread_span;
    c->dense_scan = dense_scan;
    const int rc = research_all(c);
    if (rc != GODAR_OK) { c->skin_frac = s0; c->thread_span = t0; c->dense_scan = d0; }
    return rc;
}

int godar_get_search(const godar_ctx* c, double* skin_frac, int32_t* thread_span, int32_t* dense_scan) {
    if (!c) return GODAR_ERR_ARG;
    if (skin_frac) *skin_frac = c->skin_frac;
    if (thread_span) *thread_span = c->thread_span;
    if (dense_scan) *dense_scan = c->dense_scan;
    return GODAR_OK;
}

int godar_set_repair(godar_ctx* c, int32_t lanes, int32_t pending) {
    if (!c || lanes < 1 || lanes > 64 || pending < 1 || pending > 1024) return GODAR_ERR_ARG;
    c->repair_lanes = lanes;
    c->repair_cap = pending;
    return GODAR_OK;
}

int godar_set_stage_block(godar_ctx* c, int32_t stage, int32_t threads) {
    if (!c || stage < 0 || stage >= GODAR_NSTAGES || !tunable_stage(stage) || threads < 64 ||
        threads > FLOE_BLOCK || (threads & 63))
        return GODAR_ERR_ARG;
    c->sblock[stage] = threads;
    return GODAR_OK;
}

int godar_get_stage_blocks(const godar_ctx* c, int32_t* threads) {
    if (!c || !threads) return GODAR_ERR_ARG;
    for (int k = 0; k < GODAR_NSTAGES; ++k) threads[k] = c->sblock[k];
    return GODAR_OK;
}

// Steps between superset builds the fastest floe allows at skin fraction
// `frac` (a build follows once a floe has moved 0.45 skin from where it was
// at the last one), from the floes' current speeds; at least 1, at most 5,000.
int build_period(godar_ctx* c, double frac, double* period) {
    const int N = c->N_used;
    std::vector<KinRec> kin(std::max(N, 1));
    std::vector<GeoRec> geo(std::max(N, 1));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(kin.data(), c->d.kin, sizeof(KinRec) * N, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(geo.data(), c->d.geo, sizeof(GeoRec) * N, hipMemcpyDeviceToHost));
    std::vector<double> vmax(c->Mcap, 0.0);
    for (int k = 0; k < N; ++k) {
        const int m = geo[k].member;
        if (m < 0 || m >= c->Mcap) continue;
        vmax[m] = std::max(vmax[m], std::sqrt(kin[k].u * kin[k].u + kin[k].v * kin[k].v));
    }
    double p = 5000.0;
    for (int m = 0; m < c->Mcap; ++m) {
        if (!c->members[m].active || !(vmax[m] > 0.0)) continue;
        const double skin = frac * (c->rmax[m] + c->dparams[m].rtree);
        p = std::min(p, 0.45 * skin / (vmax[m] * c->dparams[m].dt));
    }
    *period = std::max(1.0, p);
    return GODAR_OK;
}

// Search shape: the superset skin (so the cell size), the thread search's
// cell span and the dense-search threshold, one after the other.  Each
// candidate is priced as a build step plus `4 trial` steps after it: the
// mean step cost over a build and the steps until the next one,
// t_step + (t_build - t_step) / period.  The settings stay where they were
// if no candidate is cheaper.
int tune_search(godar_ctx* c, int trial) {
    const int n = std::max(4, 4 * trial);
    auto stage_sum = [&]() {
        double t = 0.0;
        for (int st = 0; st < GODAR_NSTAGES; ++st) t += c->stage_ms[st];
        return t;
    };
    auto reset = [&]() {
        for (int st = 0; st < GODAR_NSTAGES; ++st) { c->stage_ms[st] = 0.0; c->stage_n[st] = 0; }
    };
    auto price = [&](double frac, int span, int dscan, double& cost) -> int {
        int rc = godar_set_search(c, frac, span, dscan);
        if (rc != GODAR_OK) return rc;
        rc = godar_profile(c, 1);
        if (rc == GODAR_OK) { reset(); rc = godar_step(c, 1); }   // the build step
        if (rc == GODAR_OK) rc = godar_sync(c);
        if (rc != GODAR_OK) return rc;
        collect_events(c);
        const double tb = stage_sum();
        reset();
        rc = godar_step(c, n);
        if (rc == GODAR_OK) rc = godar_sync(c);
        if (rc != GODAR_OK) return rc;
        collect_events(c);
        const double ts = stage_sum() / n;
        double per = 1.0;
        rc = build_period(c, frac, &per);
        if (rc != GODAR_OK) return rc;
        cost = ts + std::max(0.0, tb - ts) / per;
        return GODAR_OK;
    };
    double best_frac = c->skin_frac;
    int best_span = c->thread_span, best_dscan = c->dense_scan;
    double best = 1e300;
    int rc = price(best_frac, best_span, best_dscan, best);
    if (rc != GODAR_OK) return rc;
    for (double f : {0.05, 0.1, 0.2}) {
        if (f == best_frac) continue;
        double cost;
        rc = price(f, best_span, best_dscan, cost);
        if (rc != GODAR_OK) return rc;
        if (cost < best) { best = cost; best_frac = f; }
    }
    for (int sp : {1, 2}) {
        if (sp == best_span) continue;
        double cost;
        rc = price(best_frac, sp, best_dscan, cost);
        if (rc != GODAR_OK) return rc;
        if (cost < best) { best = cost; best_span = sp; }
    }
    for (int ds : {48, 96, 192}) {
        if (ds == best_dscan) continue;
        double cost;
        rc = price(best_frac, best_span, ds, cost);
        if (rc != GODAR_OK) return rc;
        if (cost < best) { best = cost; best_dscan = ds; }
    }
    c->tuned_cost_ms = best;
    return godar_set_search(c, best_frac, best_span, best_dscan);
}

// Coordinate search over the block size of each tunable stage: every
// candidate runs `trial` real steps of the current members (they advance;
// results do not depend on block sizes) with stage timing, and the fastest
// is kept.  Capacity growth is handled by godar_step.
int godar_autotune(godar_ctx* c, int32_t trial) {
    if (!c || trial < 1) return GODAR_ERR_ARG;
    static const int cand[] = {64, 128, 256};
    const bool prof0 = c->prof;
    int rc = godar_step(c, 1);   // settle capacities first
    if (rc != GODAR_OK) return rc;
    for (int st = 0; st < GODAR_NSTAGES; ++st) {
        if (!tunable_stage(st)) continue;
        // The bond and wall kernels run beside the contact kernel (the
        // contacts and walls on a side stream), where single-wavefront blocks
        // fit in the room the others leave;
        // timed one kernel after another, as here, larger blocks look as
        // fast, but overlapped 256-thread bond blocks cost about 3 % of a
        // step at the mid-run of the 512-field ensemble.
        if (st == 5 || st == 7) {
            c->sblock[st] = 64;
            continue;
        }
        int best = c->sblock[st];
        double best_ms = 1e300;
        for (int b : cand) {
            c->sblock[st] = b;
            rc = godar_profile(c, 1);
            if (rc == GODAR_OK) rc = godar_step(c, trial);
            if (rc == GODAR_OK) rc = godar_sync(c);
            if (rc != GODAR_OK) { c->prof = prof0; return rc; }
            collect_events(c);
            double ms = c->stage_n[st] > 0 ? c->stage_ms[st] / (double)c->stage_n[st] : 1e300;
            if (ms < best_ms) { best_ms = ms; best = b; }
        }
        c->sblock[st] = best;
    }
    rc = tune_search(c, trial);
    if (rc != GODAR_OK) { c->prof = prof0; return rc; }
    int r2 = godar_profile(c, prof0 ? 1 : 0);
    if (r2 != GODAR_OK || prof0 || !c->use_graphs) return r2;
    // launch mode: graph replay or one-by-one launches, whichever steps
    // faster; each is timed twice, alternating, over max(32, 16 trial)
    // steps, and its faster time counts (a short trial is at the mercy of
    // a single slow step)
    const int n = std::max(32, 16 * trial);
    double t[2] = {1e300, 1e300};
    for (int rep = 0; rep < 4; ++rep) {
        const int mode = rep & 1;
        c->use_graphs = mode == 0;
        rc = godar_step(c, 1);                     // capture / settle
        if (rc != GODAR_OK) { c->use_graphs = true; return rc; }
        const auto t0 = std::chrono::steady_clock::now();
        rc = godar_step(c, n);
        if (rc == GODAR_OK) rc = godar_sync(c);
        if (rc != GODAR_OK) { c->use_graphs = true; return rc; }
        t[mode] = std::min(t[mode], std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    c->use_graphs = t[0] <= t[1];
    if (!c->use_graphs) drop_graphs(c);
    return GODAR_OK;
}

int godar_record_forces(godar_ctx* c, int32_t on) {
    if (!c) return GODAR_ERR_ARG;
    c->record = on ? 1 : 0;
    return GODAR_OK;
}

int godar_step_async(godar_ctx* c, int32_t nsteps) {
    if (!c || nsteps < 0) return GODAR_ERR_ARG;
    if (!c->pending) {
        int rc = steps_done(c, &c->pend_done0);
        if (rc != GODAR_OK) return rc;
        c->pend_parity = c->parity;
        c->pending = true;
        c->pend_launched = 0;
    }
    const bool graphs = c->use_graphs && !c->prof;
    for (int k = 0; k < nsteps; ++k) {
        // the dense-search list is laid out by the candidate search's XCD
        // shares: a new block size or floe count needs a build
        const int32_t q = qseg_of(c);
        if (q != c->built_qseg) {
            force_rebuild(c);
            c->built_qseg = q;
        }
        const bool lean = !c->need_full && c->full_hold == 0 && c->lean_run < LEAN_RUN;
        int rc = graphs ? graph_step(c, lean) : launch_step(c, lean);
        if (rc != GODAR_OK) return rc;
        c->need_full = false;
        c->lean_run = lean ? c->lean_run + 1 : 0;
        if (c->full_hold > 0 && !lean) --c->full_hold;
        ++c->pend_launched;
    }
    return GODAR_OK;
}

int godar_graphs_on(const godar_ctx* c, int32_t* on) {
    if (!c || !on) return GODAR_ERR_ARG;
    *on = c->use_graphs ? 1 : 0;
    return GODAR_OK;
}

int godar_use_graphs(godar_ctx* c, int32_t on) {
    if (!c) return GODAR_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->use_graphs = on != 0;
    if (!c->use_graphs) drop_graphs(c);
    return GODAR_OK;
}

int godar_sync(godar_ctx* c) {
    if (!c) return GODAR_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    collect_events(c);
    uint32_t e = 0;
    HIPCHK(c, hipMemcpy(&e, c->d.err, sizeof(uint32_t), hipMemcpyDeviceToHost));
    int rc = download_members(c);
    if (rc != GODAR_OK) return rc;
    bool was_pending = c->pending;
    c->pending = false;
    if (e == ERR_NEED_BUILD && was_pending) {
        // a lean step found a build due: it and the steps after it did
        // nothing; back to the last complete step, and the rest again with
        // the build kernels
        int64_t now = 0;
        rc = steps_done(c, &now);
        if (rc != GODAR_OK) return rc;
        c->parity = c->pend_parity ^ (int)((now - c->pend_done0) & 1);
        bind_parity(c);
        HIPCHK(c, hipMemset(c->d.err, 0, sizeof(uint32_t)));
        const int64_t left = c->pend_launched - (now - c->pend_done0);
        c->need_full = true;
        c->full_hold = FULL_HOLD;
        if (left <= 0) return GODAR_OK;
        rc = godar_step_async(c, (int32_t)left);
        if (rc != GODAR_OK) return rc;
        return godar_sync(c);
    }
    if (e) {
        // Roll back to the last complete step.  Steps after the failing one
        // were no-ops; the device step counters tell how many completed, which
        // fixes the pair-state parity.  Repair flags and the error word are
        // cleared so the caller can reserve more room and step again.
        if (was_pending) {
            int64_t now = 0;
            rc = steps_done(c, &now);
            if (rc != GODAR_OK) return rc;
            c->parity = c->pend_parity ^ (int)((now - c->pend_done0) & 1);
            bind_parity(c);
        }
        HIPCHK(c, hipMemset(c->d.dirty, 0, c->S));
        HIPCHK(c, hipMemset(c->d.err, 0, sizeof(uint32_t)));
        c->last_err_bits = e;
        return fail(c, GODAR_ERR_CAPACITY, "step aborted:" + err_text(e));
    }
    return GODAR_OK;
}

// Raise the capacities an aborted step ran out of (lists double, repair
// pools grow by half).  Returns false when a limit is reached.
static bool grow_for(godar_ctx* c, uint32_t bits) {
    int nCO = c->CO, nCI = c->CI, nHC = c->HC, nBC = c->BC;
    if (bits & ERR_OUTER_CAP) nCO = std::min(2 * c->CO, 0xFFFE);
    if (bits & ERR_INNER_CAP) nCI = std::min(2 * c->CI, 1 << 15);
    if (bits & ERR_HIST_CAP) nHC = std::min(2 * c->HC, 1 << 12);
    if (bits & ERR_BOND_CAP) nBC = std::min(2 * c->BC, 1 << 12);
    bool grew = nCO != c->CO || nCI != c->CI || nHC != c->HC || nBC != c->BC;
    if (grew && godar_reserve(c, nCO, nCI, nHC, nBC) != GODAR_OK) return false;
    if (bits & (ERR_REPAIR_CAP | ERR_SEED_CAP)) {
        if (c->HEAPK >= 1024) return false;
        if (hipStreamSynchronize(c->stream) != hipSuccess) return false;
        // the larger pools first: on failure the current ones stay
        uint64_t *seeds = nullptr, *heap = nullptr;
        EventOut* etab = nullptr;
        if (dalloc(&seeds, (int64_t)2 * c->SEEDK * c->S) != hipSuccess ||
            dalloc(&heap, (int64_t)2 * c->HEAPK * c->S) != hipSuccess ||
            dalloc(&etab, (int64_t)2 * c->TABK * c->S) != hipSuccess) {
            dfree(seeds); dfree(heap); dfree(etab);
            (void)hipGetLastError();
            return false;
        }
        dfree(c->d.seeds); dfree(c->d.heap); dfree(c->d.etab);
        c->d.seeds = seeds; c->d.heap = heap; c->d.etab = etab;
        c->SEEDK *= 2; c->HEAPK *= 2; c->TABK *= 2;
        grew = true;
    }
    if (bits & ERR_SUP_CAP) {
        if (c->sup_cap >= ((int64_t)1 << 34)) return false;
        if (hipStreamSynchronize(c->stream) != hipSuccess) return false;
        int64_t* sup = nullptr;
        if (dalloc(&sup, 2 * c->sup_cap) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        dfree(c->d.sup);
        c->d.sup = sup;
        c->sup_cap *= 2;
        grew = true;
    }
    if (grew) force_rebuild(c);
    return grew;
}

int godar_step(godar_ctx* c, int32_t nsteps) {
    if (!c || nsteps < 0) return GODAR_ERR_ARG;
    int64_t start = 0, now = 0;
    int rc = steps_done(c, &start);
    if (rc != GODAR_OK) return rc;
    int32_t left = nsteps;
    for (int attempt = 0; attempt < 64; ++attempt) {
        rc = godar_step_async(c, left);
        if (rc != GODAR_OK) return rc;
        rc = godar_sync(c);
        if (rc != GODAR_ERR_CAPACITY || !c->auto_grow) return rc;
        // some steps completed before the overflow: only the rest is retried
        rc = steps_done(c, &now);
        if (rc != GODAR_OK) return rc;
        left = nsteps - (int32_t)(now - start);
        std::string why = c->err;
        if (!grow_for(c, c->last_err_bits))
            return fail(c, GODAR_ERR_CAPACITY, why + " (capacity limit reached)");
        if (left <= 0) return GODAR_OK;
    }
    return fail(c, GODAR_ERR_CAPACITY, "capacity growth did not converge");
}

int godar_member_run_until(godar_ctx* c, int32_t m, int32_t tlast) {
    if (!valid_member(c, m) || tlast < 0) return fail(c, GODAR_ERR_ARG, "member_run_until: bad argument");
    int rc = download_members(c);   // keep device-side step counters
    if (rc != GODAR_OK) return rc;
    c->members[m].tstop = tlast > 0 ? tlast + 1 : 0;
    rc = upload_members_only(c);
    if (rc != GODAR_OK) return rc;
    force_rebuild(c);   // a member that runs again needs its floes back in the cell order
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GODAR_OK;
}

int godar_member_tsteps(godar_ctx* c, int32_t* tsteps) {
    if (!c || !tsteps) return GODAR_ERR_ARG;
    int rc = download_members(c);
    if (rc != GODAR_OK) return rc;
    for (int k = 0; k < c->Mcap; ++k) tsteps[k] = c->members[k].active ? c->members[k].tstep : -1;
    return GODAR_OK;
}

int godar_member_tstep(const godar_ctx* c, int32_t m, int32_t* t) {
    if (!valid_member(c, m) || !t) return GODAR_ERR_ARG;
    godar_ctx* cc = const_cast<godar_ctx*>(c);
    int rc = download_members(cc);
    if (rc != GODAR_OK) return rc;
    *t = c->members[m].tstep;
    return GODAR_OK;
}

int godar_get_floes(godar_ctx* c, int32_t m, double* x, double* y, double* u, double* v,
                    double* theta, double* omega, double* r, double* h, double* mass,
                    double* inertia, double* hfa, double* hfw) {
    if (!valid_member(c, m)) return fail(c, GODAR_ERR_ARG, "get_floes: bad member");
    spill(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const Member& mm = c->members[m];
    std::vector<FloeRec> recs(mm.n);
    HIPCHK(c, hipMemcpy(recs.data(), c->d.rec + mm.first, sizeof(FloeRec) * mm.n, hipMemcpyDeviceToHost));
    double* dst[] = {x, y, u, v, theta, omega, r, h, mass, inertia, hfa, hfw};
    const std::vector<int32_t>& pm = c->perm[m];   // storage offset -> caller's index
    for (int t = 0; t < mm.n; ++t) {
        const FloeRec& q = recs[t];
        double val[12] = {q.x, q.y, q.u, q.v, q.theta, q.omega, q.r, q.h, q.mass, q.inertia, q.hfa, q.hfw};
        for (int k = 0; k < 12; ++k)
            if (dst[k]) dst[k][pm[t]] = val[k];
    }
    return GODAR_OK;
}

int godar_get_forces(godar_ctx* c, int32_t m, double* f) {
    if (!valid_member(c, m) || !f) return fail(c, GODAR_ERR_ARG, "get_forces: bad argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const Member& mm = c->members[m];
    const std::vector<int32_t>& pm = c->perm[m];
    std::vector<double> row(mm.n);
    for (int k = 0; k < 33; ++k) {
        HIPCHK(c, hipMemcpy(row.data(), c->d.fout + k * c->S + mm.first, sizeof(double) * mm.n,
                            hipMemcpyDeviceToHost));
        for (int t = 0; t < mm.n; ++t) f[(int64_t)k * mm.n + pm[t]] = row[t];
    }
    return GODAR_OK;
}

int godar_get_walls(godar_ctx* c, int32_t m, double* w) {
    if (!valid_member(c, m) || !w) return fail(c, GODAR_ERR_ARG, "get_walls: bad argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const Member& mm = c->members[m];
    const std::vector<int32_t>& pm = c->perm[m];
    std::vector<double> row(mm.n);
    for (int k = 0; k < 4; ++k) {
        HIPCHK(c, hipMemcpy(row.data(), c->d.wbc + k * c->S + mm.first, sizeof(double) * mm.n,
                            hipMemcpyDeviceToHost));
        for (int t = 0; t < mm.n; ++t) w[(int64_t)k * mm.n + pm[t]] = row[t];
    }
    return GODAR_OK;
}

int godar_get_pairs(godar_ctx* c, int32_t m, int32_t kind, int32_t cap, int32_t* pi, int32_t* pj,
                    int32_t* count) {
    if (!valid_member(c, m) || !count || cap < 0 || (cap > 0 && (!pi || !pj)))
        return fail(c, GODAR_ERR_ARG, "get_pairs: bad argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const Member& mm = c->members[m];
    int64_t S = c->S;
    std::vector<std::pair<int32_t, int32_t>> out;
    if (kind == GODAR_PAIRS_CANDIDATE || kind == GODAR_PAIRS_CONTACT) {
        // the member's floes occupy a contiguous range of this step's cell order
        std::vector<int32_t> pos(mm.n);
        HIPCHK(c, hipMemcpy(pos.data(), c->d.pos_of + mm.first, sizeof(int32_t) * mm.n, hipMemcpyDeviceToHost));
        int32_t p0 = *std::min_element(pos.begin(), pos.end());
        int32_t len = *std::max_element(pos.begin(), pos.end()) - p0 + 1;
        std::vector<int32_t> cnt(len);
        HIPCHK(c, hipMemcpy(cnt.data(), c->d.ocnt + p0, sizeof(int32_t) * len, hipMemcpyDeviceToHost));
        int32_t cmax = 0;
        for (int32_t v : cnt) cmax = std::max(cmax, v);
        cmax = std::min(cmax, c->CO);
        std::vector<int64_t> oj((int64_t)cmax * len);
        std::vector<uint32_t> ef((int64_t)cmax * len);
        for (int s = 0; s < cmax; ++s) {
            HIPCHK(c, hipMemcpy(oj.data() + (int64_t)s * len, c->d.oj + s * S + p0,
                                sizeof(int64_t) * len, hipMemcpyDeviceToHost));
            HIPCHK(c, hipMemcpy(ef.data() + (int64_t)s * len, c->d.evf + s * S + p0,
                                sizeof(uint32_t) * len, hipMemcpyDeviceToHost));
        }
        for (int i = 0; i < mm.n; ++i) {
            int32_t r0 = pos[i] - p0;
            for (int s = 0; s < cnt[r0] && s < cmax; ++s) {
                uint32_t fl = ef[(int64_t)s * len + r0];
                if (fl & EV_REMOVED) continue;
                if (kind == GODAR_PAIRS_CONTACT && !(fl & EV_CONTACT)) continue;
                out.push_back({c->perm[m][i], (int32_t)(oj[(int64_t)s * len + r0] >> 32) - mm.first});
            }
        }
    } else if (kind == GODAR_PAIRS_BOND) {
        int o = c->parity;
        std::vector<int32_t> cnt(mm.n);
        HIPCHK(c, hipMemcpy(cnt.data(), c->bcnt[o] + mm.first, sizeof(int32_t) * mm.n, hipMemcpyDeviceToHost));
        std::vector<BondRec> rec(mm.n);
        for (int q = 0; q < c->BC; ++q) {
            HIPCHK(c, hipMemcpy(rec.data(), c->bond[o] + q * S + mm.first, sizeof(BondRec) * mm.n,
                                hipMemcpyDeviceToHost));
            for (int i = 0; i < mm.n; ++i)
                if (q < cnt[i] && rec[i].j >= 0 && rec[i].bond == 1)
                    out.push_back({c->perm[m][i], rec[i].j - mm.first});
        }
    } else {
        return fail(c, GODAR_ERR_ARG, "get_pairs: bad kind");
    }
    std::sort(out.begin(), out.end());
    *count = (int32_t)out.size();
    for (int32_t k = 0; k < cap && k < (int32_t)out.size(); ++k) { pi[k] = out[k].first; pj[k] = out[k].second; }
    return GODAR_OK;
}

int godar_get_bond_state(godar_ctx* c, int32_t m, int32_t cap, double* st, int32_t* count) {
    if (!valid_member(c, m) || !count || cap < 0 || (cap > 0 && !st))
        return fail(c, GODAR_ERR_ARG, "get_bond_state: bad argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const Member& mm = c->members[m];
    int64_t S = c->S;
    int o = c->parity;
    std::vector<int32_t> cnt(mm.n);
    HIPCHK(c, hipMemcpy(cnt.data(), c->bcnt[o] + mm.first, sizeof(int32_t) * mm.n, hipMemcpyDeviceToHost));
    std::vector<std::pair<std::pair<int32_t, int32_t>, BondRec>> out;
    std::vector<BondRec> rec(mm.n);
    for (int q = 0; q < c->BC; ++q) {
        HIPCHK(c, hipMemcpy(rec.data(), c->bond[o] + q * S + mm.first, sizeof(BondRec) * mm.n,
                            hipMemcpyDeviceToHost));
        for (int i = 0; i < mm.n; ++i)
            if (q < cnt[i] && rec[i].j >= 0 && rec[i].bond == 1)
                out.push_back({{c->perm[m][i], rec[i].j - mm.first}, rec[i]});
    }
    std::sort(out.begin(), out.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    *count = (int32_t)out.size();
    for (int32_t k = 0; k < cap && k < (int32_t)out.size(); ++k) {
        const BondRec& b = out[k].second;
        double v[11] = {(double)b.bond, b.damage, b.deltanb, b.deltatb, b.trb_ji, b.trb_ij, b.rb, b.hb, b.lb, b.sb, b.ib};
        std::memcpy(st + 11 * (int64_t)k, v, sizeof(v));
    }
    return GODAR_OK;
}

int godar_get_totals(godar_ctx* c, int64_t* out) {
    if (!c || !out) return GODAR_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t* dout = nullptr;
    HIPCHK(c, hipMalloc(&dout, 5 * sizeof(int64_t)));
    HIPCHK(c, hipMemset(dout, 0, 5 * sizeof(int64_t)));
    sync_ds(c);
    if (c->N_used > 0)
        hipLaunchKernelGGL(k_totals, dim3((c->N_used + 255) / 256), dim3(256), 0, c->stream, c->d, dout);
    hipLaunchKernelGGL(k_member_totals, dim3(1), dim3(256), 0, c->stream, c->d, c->Mcap, dout + 3);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(out, dout, 3 * sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out + 6, dout + 3, 2 * sizeof(int64_t), hipMemcpyDeviceToHost);
    std::vector<int32_t> q((size_t)NQUEUE * XCD_COUNT * QN_STRIDE, 0);
    if (e == hipSuccess)
        e = hipMemcpy(q.data(), c->d.qn, q.size() * sizeof(int32_t), hipMemcpyDeviceToHost);
    for (int k = 0; k < 3; ++k) {
        out[3 + k] = 0;
        for (int x = 0; x < XCD_COUNT; ++x) out[3 + k] += q[(size_t)(k * XCD_COUNT + x) * QN_STRIDE];
    }
    (void)hipFree(dout);
    HIPCHK(c, e);
    return GODAR_OK;
}

// godar_get_totals without allocating: the same kernel sums the pair totals
// into this GPU's `scratch` (through its device address), a second kernel the
// queue totals; both are cleared by a kernel first (no memset node), and one
// copy brings all six back.
int godar_get_totals_noalloc(godar_ctx* c, int64_t* out) {
    if (!c || !out) return GODAR_ERR_ARG;
    if (!c->totals_scratch) return fail(c, GODAR_ERR_STATE, "get_totals_noalloc: no scratch address");
    int cur = -1;
    HIPCHK(c, hipGetDevice(&cur));
    if (cur != c->device)
        return fail(c, GODAR_ERR_STATE, "get_totals_noalloc: call with the context's GPU current (" +
                                            std::to_string(c->device) + ")");
    int64_t* w = c->totals_scratch;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    sync_ds(c);
    hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, c->stream, w, GODAR_NTOTALS);
    if (c->N_used > 0)
        hipLaunchKernelGGL(k_totals, dim3((c->N_used + 255) / 256), dim3(256), 0, c->stream, c->d, w);
    hipLaunchKernelGGL(k_queue_totals, dim3(1), dim3(64), 0, c->stream, c->d.qn, w + 3);
    hipLaunchKernelGGL(k_member_totals, dim3(1), dim3(256), 0, c->stream, c->d, c->Mcap, w + 6);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, w, GODAR_NTOTALS * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GODAR_OK;
}

static const char* kStageNames[GODAR_NSTAGES] = {
    "bin", "sort", "cells", "candidates", "contacts", "bonds", "unused", "walls", "repair",
    "fixup", "stress", "integrate"};

int godar_profile(godar_ctx* c, int32_t on) {
    if (!c) return GODAR_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    collect_events(c);
    c->prof = on != 0;
    for (int k = 0; k < GODAR_NSTAGES; ++k) { c->stage_ms[k] = 0; c->stage_n[k] = 0; }
    return GODAR_OK;
}

int godar_get_profile(godar_ctx* c, double* ms, int64_t* launches) {
    if (!c || !ms || !launches) return GODAR_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    collect_events(c);
    for (int k = 0; k < GODAR_NSTAGES; ++k) { ms[k] = c->stage_ms[k]; launches[k] = c->stage_n[k]; }
    return GODAR_OK;
}

const char* godar_stage_name(int32_t k) {
    return (k >= 0 && k < GODAR_NSTAGES) ? kStageNames[k] : "";
}

int godar_get_stats(godar_ctx* c, int32_t m, int64_t* stats) {
    if (!valid_member(c, m) || !stats) return fail(c, GODAR_ERR_ARG, "get_stats: bad argument");
    int32_t cnt = 0;
    int rc = godar_get_pairs(c, m, GODAR_PAIRS_CANDIDATE, 0, nullptr, nullptr, &cnt);
    if (rc != GODAR_OK) return rc;
    stats[0] = cnt;
    rc = godar_get_pairs(c, m, GODAR_PAIRS_CONTACT, 0, nullptr, nullptr, &cnt);
    if (rc != GODAR_OK) return rc;
    stats[1] = cnt;
    int64_t dev[4];
    HIPCHK(c, hipMemcpy(dev, c->d.stats + 4 * m, sizeof(dev), hipMemcpyDeviceToHost));
    stats[2] = dev[2];
    stats[3] = dev[3];
    return GODAR_OK;
}

}  // extern "C"
