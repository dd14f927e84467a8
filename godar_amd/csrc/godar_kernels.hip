// Kernels of the GODAR step pipeline (one launch sequence per step):
//
//   k_begin -> [build steps: radix sort -> k_ranges -> k_sup ...] -> k_cand
//    -> (k_contact -> k_walls on a side stream) beside k_bond
//    -> k_comp -> k_repair -> k_fixq -> k_fixup
//    -> k_final (which also gathers the bond terms)
//
// Sequential semantics of the reference (stepper.f90:71-280 visits floes i
// in index order and, for each, its candidates j > i in ascending distance;
// ridging reshapes floes in place, so later pairs see the new radius and
// thickness) are reproduced as follows:
//   * The event kernels evaluate every pair event with the start-of-step
//     floe properties.  Contact, bond and sheltering terms feed disjoint
//     accumulators, so each kind has a kernel of its own: the candidate
//     search k_cand also computes every floe's sheltering minima and queues
//     the floes that have contact, bond or wall work for k_contact, k_bond
//     and k_walls.  One thread per
//     floe gathers its own contributions in the reference's per-floe order:
//     first the pairs where it is the inner index (ascending outer index),
//     then its own candidates in distance order.  Contact and sheltering
//     terms are evaluated by both floes of a pair; a bond is evaluated once,
//     by its owner, which stores the terms; k_final gathers them.  No atomics,
//     and FP64 sums stay identical to the reference.
//   * An event's speculative result is exact unless one of its floes was
//     reshaped earlier in the sequence.  Events that ridge are seeds; the
//     per-member k_repair walks, in sequence order, the seeds and every later
//     event of each reshaped floe, recomputes them with the current floe
//     properties and records the results.
//   * k_fixq lists the floes touched by the repair, k_fixup re-gathers them,
//     k_final applies the final floe properties, forcing, Coriolis, totals
//     and integration.
//
// Layout: the per-step kernels run over floes in cell order ("pos"), reading
// a sorted copy of the 96-byte floe records, so that a floe's neighbours sit
// close in memory.  Pair state persists per floe id.  List entries pack the
// partner as (id << 32 | pos).
#include "godar_engine.h"
#include "godar_physics.h"
#include <type_traits>

namespace godar {

__device__ __forceinline__ void set_err(const DS& S, uint32_t bit) { atomicOr(S.err, bit); }

// LDS written by some lanes of a one-wavefront block, then read by others:
// the wavefront's LDS operations complete in order, so waiting for its own
// outstanding ones (and keeping the compiler from moving accesses across)
// is enough; no block barrier, which would also wait for global stores.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0); vmcnt, expcnt unchanged (gfx9 encoding)
    __builtin_amdgcn_wave_barrier();
}

// Workgroups are dispatched to the 8 XCDs round-robin, and each XCD has its
// own L2.  Kernels that read their neighbours' rows take this logical block
// index instead of blockIdx.x, so that each XCD works on one contiguous eighth
// of the floes (whose neighbours then sit in the same L2).  The grid must be a
// multiple of XCD_COUNT blocks (launch_grid on the host).
__device__ __forceinline__ int xcd_block() {
    const int per = (int)(gridDim.x / XCD_COUNT);
    return (int)(blockIdx.x % XCD_COUNT) * per + (int)(blockIdx.x / XCD_COUNT);
}

// Entry of a work queue of length qn for this thread, or -1.  Each XCD takes
// one contiguous share of the queue (the grid, sized for all floes, always
// has enough threads per XCD), so that short queues still spread over all
// XCDs and each XCD's share sits together in its L2.
__device__ __forceinline__ int queue_slot(int qn) {
    const int per = (qn + XCD_COUNT - 1) / XCD_COUNT;
    const int local = (int)(blockIdx.x / XCD_COUNT) * (int)blockDim.x + (int)threadIdx.x;
    if (local >= per) return -1;
    const int t = (int)(blockIdx.x % XCD_COUNT) * per + local;
    return t < qn ? t : -1;
}

__device__ __forceinline__ int32_t* qcount(const DS& S, int c, int x) {
    return S.qn + (c * XCD_COUNT + x) * QN_STRIDE;
}

// Queue c read as one sequence (its segments in XCD order): the segment
// counts and the total.
struct QueueView {
    int cnt[XCD_COUNT];
    int total;
};
__device__ __forceinline__ QueueView queue_view(const DS& S, int c) {
    QueueView v;
    v.total = 0;
#pragma unroll
    for (int x = 0; x < XCD_COUNT; ++x) {
        v.cnt[x] = *qcount(S, c, x);
        v.total += v.cnt[x];
    }
    return v;
}
// position held by entry t (0 <= t < total) of the sequence
__device__ __forceinline__ int queue_at(const DS& S, int c, const QueueView& v, int t) {
    int idx = 0;
#pragma unroll
    for (int x = 0; x < XCD_COUNT; ++x) {
        idx = (t >= 0 && t < v.cnt[x]) ? x * S.qseg + t : idx;
        t -= v.cnt[x];
    }
    return S.qlist[(int64_t)c * S.S + idx];
}
// Slot of this thread's it-th entry of a queue of length qn, or -1: the
// grid may be smaller than the queue (the host caps the queue kernels' grids
// at QUEUE_GRID_WAVES wavefronts), so each thread takes every
// (threads per XCD)-th entry of its XCD's share.
__device__ __forceinline__ int queue_slot_it(int qn, int it) {
    const int per = (qn + XCD_COUNT - 1) / XCD_COUNT;
    const int local = ((int)(blockIdx.x / XCD_COUNT) + it * (int)(gridDim.x / XCD_COUNT)) * (int)blockDim.x +
                      (int)threadIdx.x;
    if (local >= per) return -1;
    const int t = (int)(blockIdx.x % XCD_COUNT) * per + local;
    return t < qn ? t : -1;
}
// the queue entry of this thread (split over the XCDs as queue_slot), or -1
__device__ __forceinline__ int queue_entry(const DS& S, int c) {
    const QueueView v = queue_view(S, c);
    const int t = queue_slot(v.total);
    return t < 0 ? -1 : queue_at(S, c, v, t);
}
// appends position k to queue c from a single lane
__device__ __forceinline__ void enqueue_one(const DS& S, int c, int k) {
    const int x = k / S.qseg;
    S.qlist[(int64_t)c * S.S + x * S.qseg + atomicAdd(qcount(S, c, x), 1)] = k;
}

__device__ __forceinline__ int ent_id(int64_t e) { return (int)(e >> 32); }
__device__ __forceinline__ int ent_pos(int64_t e) { return (int)(uint32_t)(e & 0xffffffffll); }
__device__ __forceinline__ int64_t ent(int id, int pos) {
    return ((int64_t)id << 32) | (int64_t)(uint32_t)pos;
}

// floe at position k of this step's cell order
__device__ __forceinline__ Floe floe_at(const DS& S, int k) {
    const GeoRec g = S.geo[k];
    const ShelRec h = S.shel[k];
    const KinRec q = S.kin[k];
    Floe F;
    F.x = g.x; F.y = g.y; F.u = q.u; F.v = q.v; F.omega = q.omega;
    F.r = g.r; F.h = q.h; F.mass = q.mass; F.inertia = q.inertia; F.hfa = h.hfa; F.hfw = h.hfw;
    return F;
}

__device__ __forceinline__ double2 wind_at(const DS& S, int k) {
    const ShelRec h = S.shel[k];
    return make_double2(h.wx, h.wy);
}

// Sets flag bits of position k's wflag byte from any lane or wavefront: the
// byte shares a 32-bit word with three neighbours whose bits others may be
// setting at the same time (the array is allocated in whole words).
__device__ __forceinline__ void wflag_set(const DS& S, int k, uint32_t bits) {
    atomicOr(reinterpret_cast<uint32_t*>(S.wflag + (k & ~3)), bits << (8 * (k & 3)));
}

// Contact components of the step (union-find over the speculative contacts,
// at positions): k_cand makes every floe its own root, k_contact unites the
// floes of each contact, hooking the larger root under the smaller with a
// compare-and-swap (a root's parent changes only there), and k_comp points
// every contact floe at its root.  Parents are read and written at device
// scope: other wavefronts change them during k_contact.
__device__ __forceinline__ int comp_parent(const int32_t* p, int x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int comp_find(int32_t* p, int x) {
    for (;;) {
        const int px = comp_parent(p, x);
        if (px == x) return x;
        const int ppx = comp_parent(p, px);
        // path halving: any ancestor may stand in for the parent
        if (ppx != px) __hip_atomic_store(p + x, ppx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = px;
    }
}
__device__ __forceinline__ void comp_unite(int32_t* p, int a, int b) {
    for (;;) {
        a = comp_find(p, a);
        b = comp_find(p, b);
        if (a == b) return;
        if (a < b) { const int t = a; a = b; b = t; }
        if (atomicCAS(p + a, a, b) == a) return;
    }
}

// floe at storage slot f and position k with the properties current during
// the ordered repair
__device__ __forceinline__ Floe load_floe_cur(const DS& S, int f, int k) {
    // the current rows are read with the records (one round of loads) and
    // taken where the floe was reshaped
    Floe F = floe_at(S, k);
    const bool d = (S.dirty[f] & DIRTY_SHAPE) != 0;
    const double r = S.cur[0 * S.S + f], h = S.cur[1 * S.S + f], mass = S.cur[2 * S.S + f];
    const double inertia = S.cur[3 * S.S + f], hfa = S.cur[4 * S.S + f], hfw = S.cur[5 * S.S + f];
    F.r = d ? r : F.r; F.h = d ? h : F.h; F.mass = d ? mass : F.mass;
    F.inertia = d ? inertia : F.inertia; F.hfa = d ? hfa : F.hfa; F.hfw = d ? hfw : F.hfw;
    return F;
}

__device__ __forceinline__ int cell_coord(double p, double cell, int g) {
    double q = floor(p / cell);
    if (!(q >= 0.0)) return 0;          // also catches NaN
    if (q >= (double)(g - 1)) return g - 1;
    return (int)q;
}

// Old pair state of (i, j): owner i by storage slot, partner j by identity.
// Lists are searched with their first LIST_PEEK entries read at once (a
// floe's lists are short): one round of loads instead of one per entry.
constexpr int LIST_PEEK = 4;

// Index of the first of owner i's n records (slot-major rows of stride SS)
// whose partner is j (j >= 0; empty records hold -1), or -1.
// The first PEEK rows are requested together (one test of n, no test per
// row; rows past n stand in for the last one, so no other line is read).
template <int PEEK = LIST_PEEK, typename Rec>
__device__ __forceinline__ int find_partner(const Rec* list, int64_t SS, int i, int n, int j) {
    int id[PEEK];
#pragma unroll
    for (int p = 0; p < PEEK; ++p) id[p] = -1;
    if (n > 0) {
#pragma unroll
        for (int p = 0; p < PEEK; ++p) id[p] = list[(int64_t)min(p, n - 1) * SS + i].j;
    }
    int q = -1;
#pragma unroll
    for (int p = PEEK - 1; p >= 0; --p) q = (p < n && id[p] == j) ? p : q;
    for (int p = PEEK; q < 0 && p < n; ++p)
        if (list[(int64_t)p * SS + i].j == j) q = p;
    return q;
}

// Slot of partner g among the first n entries of the slot-major list row kp
// (outer oj or inner ik), or -1; the first LIST_PEEK entries are requested
// together (as find_partner), the rest read one by one.
__device__ __forceinline__ int find_entry(const int64_t* list, int64_t SS, int kp, int n, int g) {
    int id[LIST_PEEK];
#pragma unroll
    for (int s = 0; s < LIST_PEEK; ++s) id[s] = -1;
    if (n > 0) {
#pragma unroll
        for (int s = 0; s < LIST_PEEK; ++s) id[s] = ent_id(list[(int64_t)min(s, n - 1) * SS + kp]);
    }
    int r = -1;
#pragma unroll
    for (int s = LIST_PEEK - 1; s >= 0; --s) r = (s < n && id[s] == g) ? s : r;
    for (int s = LIST_PEEK; r < 0 && s < n; ++s)
        if (ent_id(list[(int64_t)s * SS + kp]) == g) r = s;
    return r;
}

// (PEEK 1: one entry at a time, for kernels short of registers)
template <int PEEK = LIST_PEEK>
__device__ __forceinline__ void load_pair_state(const DS& S, int i, int j, PairState& ps) {
    ps.deltat = 0.0; ps.thetarelc = 0.0; ps.fcn = 0.0; ps.fct = 0.0;
    ps.bond = 0; ps.has_bond_rec = false;
    ps.damage = ps.deltanb = ps.deltatb = ps.trb_ji = ps.trb_ij = 0.0;
    ps.rb = ps.hb = ps.lb = ps.sb = ps.ib = 0.0;
    // both lists are searched at once
    const int hc = S.hcnt_old[i], bc = S.bcnt_old[i];
    const int hq = find_partner<PEEK>(S.hist_old, S.S, i, hc, j);
    const int bq = find_partner<PEEK>(S.bond_old, S.S, i, bc, j);
    if (hq >= 0) {
        const HistRec& hr = S.hist_old[(int64_t)hq * S.S + i];
        ps.deltat = hr.deltat; ps.thetarelc = hr.thetarelc; ps.fcn = hr.fcn; ps.fct = hr.fct;
    }
    if (bq >= 0) {
        const BondRec& b = S.bond_old[(int64_t)bq * S.S + i];
        ps.has_bond_rec = true; ps.bond = b.bond; ps.damage = b.damage;
        ps.deltanb = b.deltanb; ps.deltatb = b.deltatb; ps.trb_ji = b.trb_ji; ps.trb_ij = b.trb_ij;
        ps.rb = b.rb; ps.hb = b.hb; ps.lb = b.lb; ps.sb = b.sb; ps.ib = b.ib;
    }
}

// Record copy word by word (a struct copy is lowered through scratch memory).
static_assert(sizeof(BondRec) % 8 == 0, "BondRec is copied in 8-byte words");
__device__ __forceinline__ void copy_bond(BondRec* dst, const BondRec* src) {
    const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
    uint64_t* d = reinterpret_cast<uint64_t*>(dst);
#pragma unroll
    for (int w = 0; w < (int)(sizeof(BondRec) / 8); ++w) d[w] = s[w];
}

__device__ __forceinline__ BondRec bond_from(const PairState& ps, int j) {
    BondRec b;
    b.j = j; b.bond = ps.bond; b.damage = ps.damage; b.deltanb = ps.deltanb; b.deltatb = ps.deltatb;
    b.trb_ji = ps.trb_ji; b.trb_ij = ps.trb_ij; b.rb = ps.rb; b.hb = ps.hb; b.lb = ps.lb;
    b.sb = ps.sb; b.ib = ps.ib;
    return b;
}

// Gather of one event into floe i's accumulators (stepper.f90:113-146,
// 165-198, 215-223), operand order kept; one helper per accumulator kind.
__device__ __forceinline__ void apply_contact_i(double* a, const EventOut& o) {
    if (o.flags & EV_CONTACT) {
        a[A_FCX] = a[A_FCX] - o.fcn_c + o.fcr_s;
        a[A_FCY] = a[A_FCY] - o.fcn_s - o.fcr_c;
        a[A_MC] = a[A_MC] - o.ri_fct - o.mcc;
        a[A_TAC] = a[A_TAC] + o.ac;
        a[A_PC] = a[A_PC] - o.fcn_ac;
    }
}
template <typename E>   // EventOut or BondTerms
__device__ __forceinline__ void add_bond_i(double* a, const E& o) {
    a[A_FBX] = a[A_FBX] + o.fbn_c - o.fbt_s;
    a[A_FBY] = a[A_FBY] + o.fbn_s + o.fbt_c;
    a[A_MB] = a[A_MB] - o.lri_fbt - o.mbb_ji;
    a[A_TAB] = a[A_TAB] + o.sb;
    a[A_PB] = a[A_PB] - o.fbn_sb;
}
__device__ __forceinline__ void apply_bond_i(double* a, const EventOut& o) {
    if (o.flags & EV_BONDED) add_bond_i(a, o);
}
__device__ __forceinline__ void apply_i(double* a, const EventOut& o) {
    apply_contact_i(a, o);
    apply_bond_i(a, o);
    if (o.flags & EV_SHELTER) {
        a[A_SHA] = fmin(a[A_SHA], o.hsfa_ji);
        a[A_SHW] = fmin(a[A_SHW], o.hsfw_ji);
    }
}

// ... and into floe j's.
__device__ __forceinline__ void apply_contact_j(double* a, const EventOut& o) {
    if (o.flags & EV_CONTACT) {
        a[A_FCX] = a[A_FCX] + o.fcn_c - o.fcr_s;
        a[A_FCY] = a[A_FCY] + o.fcn_s + o.fcr_c;
        a[A_MC] = a[A_MC] - o.rj_fct + o.mcc;
        a[A_TAC] = a[A_TAC] + o.ac;
        a[A_PC] = a[A_PC] - o.fcn_ac;
    }
}
template <typename E>
__device__ __forceinline__ void add_bond_j(double* a, const E& o) {
    a[A_FBX] = a[A_FBX] - o.fbn_c + o.fbt_s;
    a[A_FBY] = a[A_FBY] - o.fbn_s - o.fbt_c;
    a[A_MB] = a[A_MB] - o.lrj_fbt - o.mbb_ij;
    a[A_TAB] = a[A_TAB] + o.sb;
    a[A_PB] = a[A_PB] - o.fbn_sb;
}
__device__ __forceinline__ void apply_bond_j(double* a, const EventOut& o) {
    if (o.flags & EV_BONDED) add_bond_j(a, o);
}
__device__ __forceinline__ void apply_j(double* a, const EventOut& o) {
    apply_contact_j(a, o);
    apply_bond_j(a, o);
    if (o.flags & EV_SHELTER) {
        a[A_SHA] = fmin(a[A_SHA], o.hsfa_ij);
        a[A_SHW] = fmin(a[A_SHW], o.hsfw_ij);
    }
}

__device__ __forceinline__ void acc_init(double* a) {
#pragma unroll
    for (int k = 0; k < NACC; ++k) a[k] = 0.0;
    a[A_SHA] = 1.0;
    a[A_SHW] = 1.0;
}

// ---------------------------------------------------------------------
// 1. cell keys (one per floe id)
// ---------------------------------------------------------------------
// rec from the ordered copies (k_final's state) where they are newer, for
// the host (`host`, which clears the flag itself) before it reads or writes
// floe records or changes the floe count; k_begin does the same at the start
// of build and refresh steps.  Positions whose copy is of no live member are
// skipped: their records are current.
__global__ void k_spill(DS S, int host) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S.N || !S.rebuild[2] || (!host && !S.rebuild[0] && !S.rebuild[1])) return;
    const GeoRec g = S.geo[k];
    const uint32_t f = S.cval[k];
    if (g.member < 0 || f >= (uint32_t)S.N) return;
    const KinRec q = S.kin[k];
    const ShelRec h = S.shel[k];
    FloeRec r;
    r.x = g.x; r.y = g.y; r.u = q.u; r.v = q.v; r.omega = q.omega;
    r.r = g.r; r.h = q.h; r.mass = q.mass; r.inertia = q.inertia; r.hfa = h.hfa; r.hfw = h.hfw;
    r.theta = q.theta;
    S.rec[f] = r;
}

// First kernel of every step, over max(N, cells) threads.  A step launched
// without the build kernels (lean) that finds a build or a refresh due sets
// ERR_NEED_BUILD and does nothing else (every later kernel of the step then
// returns at once and the host runs the step again with them); with the
// build kernels, BUILD_SOON alone is enough to build.  Then:
//  * the queue counters (the dense-search list, queue 3, only on build steps:
//    it lists positions, which change only then) and, on build steps, the
//    cell ranges;
//  * thread i as position: rec from the ordered copies where they are newer,
//    on build and refresh steps (k_spill's work; k_ranges clears the flag);
//  * thread i as storage slot: the per-member resets of the step and, on
//    build steps, the floe's cell key for the sort, written to the sort's
//    input buffers (bkey, bval).  The position is read from the ordered copy
//    where that is newer than rec (it is the same value the spill writes),
//    so the spill and the keys need no order between them.
__global__ void k_begin(DS S, int lean, int nmembers) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (*S.err) return;
    if (lean && ((S.rebuild[0] & BUILD_DUE) || S.rebuild[1])) {
        if (i == 0) atomicOr(S.err, ERR_NEED_BUILD);
        return;
    }
    const bool build = !lean && S.rebuild[0] != 0, refresh = !lean && S.rebuild[1] != 0;
    const bool newer = S.rebuild[2] != 0;
    // (the dense-search queue stands between builds)
    if (i < NQUEUE * XCD_COUNT && (build || i / XCD_COUNT != 3)) S.qn[i * QN_STRIDE] = 0;
    if (build && i <= S.ncells) S.crange[i] = make_int2(0, 0);
    // per-member resets (thread m for member m)
    if (i < nmembers && member_live(S.members[i])) {
        S.nseeds[i] = 0; S.netab[i] = 0; S.nwtab[i] = 0;
        S.stats[4 * i + 2] = 0; S.stats[4 * i + 3] = 0;
        S.tcur[i] = S.members[i].tstep;
    }
    if (lean || i >= S.N) return;   // (a lean step's grid covers the counters and members only)
    if (newer && (build || refresh)) {
        const GeoRec g = S.geo[i];
        const uint32_t fs = S.cval[i];
        if (g.member >= 0 && fs < (uint32_t)S.N) {
            const KinRec q = S.kin[i];
            const ShelRec h = S.shel[i];
            FloeRec r;
            r.x = g.x; r.y = g.y; r.u = q.u; r.v = q.v; r.omega = q.omega;
            r.r = g.r; r.h = q.h; r.mass = q.mass; r.inertia = q.inertia; r.hfa = h.hfa; r.hfw = h.hfw;
            r.theta = q.theta;
            S.rec[fs] = r;
        }
    }
    const int f = i;
    if (!build) return;           // the cell order of the last build stands
    const int m = S.member_of[f];
    const bool live = m >= 0 && member_live(S.members[m]);
    S.bval[f] = (uint32_t)f;
    // floes of free slots and of members past their stop step sort last and
    // take no part in the step
    if (!live) { S.bkey[f] = (uint32_t)S.ncells; return; }
    const DevParams& P = S.params[m];
    double x, y;
    const GeoRec g = newer ? S.geo[S.pos_of[f]] : GeoRec{};
    if (newer && g.member >= 0) { x = g.x; y = g.y; }
    else { x = S.rec[f].x; y = S.rec[f].y; }
    int cx = cell_coord(x, P.cell, P.gx), cy = cell_coord(y, P.cell, P.gy);
    S.bkey[f] = (uint32_t)(P.cell_base + cy * P.gx + cx);
}

// The floe records in the current order and the wind at each floe
// (winds_currents, forcing.f90:1-29); on build steps the cell ranges over the
// sorted keys.  Between builds the order stands and k_final writes these
// copies for the next step itself, so this runs only on build steps and after
// the copies went stale (a member stopped, the host changed the wind).
__global__ void k_ranges(DS S) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    // k_begin has brought rec up to date where this step reads it
    if (k == 0 && !*S.err && (S.rebuild[0] || S.rebuild[1])) S.rebuild[2] = 0;
    if (k >= S.N || *S.err || (!S.rebuild[0] && !S.rebuild[1])) return;
    int f = (int)S.cval[k];
    int m = S.member_of[f];
    if (m >= 0 && !member_live(S.members[m])) m = -1;
    const FloeRec fr = S.rec[f];
    GeoRec g;
    g.x = fr.x; g.y = fr.y; g.r = fr.r; g.id = S.gid_of[f]; g.member = m;
    S.geo[k] = g;
    KinRec q;
    q.u = fr.u; q.v = fr.v; q.omega = fr.omega; q.h = fr.h; q.mass = fr.mass; q.inertia = fr.inertia;
    q.theta = fr.theta;
    S.kin[k] = q;
    S.pos_of[f] = k;
    double ua = 0.0, va = 0.0;
    if (m >= 0) {
        S.pos_id[g.id] = k;
        if (*S.rebuild) {
            const uint32_t key = S.ckey[k];
            if (k == 0 || S.ckey[k - 1] != key) S.crange[key].x = k;
            if (k == S.N - 1 || S.ckey[k + 1] != key) S.crange[key].y = k + 1;
        }
        const DevParams& P = S.params[m];
        ua = P.ua; va = P.va;
        // the wind at the start of step tcur: model time (tcur - 1) dt
        if (P.wind_grid >= 0)
            sample_wind(S.wgrids[P.wind_grid], S.wpool, fr.x, fr.y, fr.r,
                        (double)(S.tcur[m] - 1) * P.dt, ua, va);
    }
    ShelRec h;
    h.hfa = fr.hfa; h.hfw = fr.hfw; h.wx = ua; h.wy = va;
    S.shel[k] = h;
}

// Appends position k to work queue c, one atomic per wavefront.  Every lane
// of the wavefront must call it; the wavefront's positions lie in one XCD
// share of the candidate search.
__device__ __forceinline__ void enqueue(const DS& S, bool want, int c, int k) {
    const uint64_t mask = __ballot(want);
    if (!mask) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    const int x = __shfl(k, leader) / S.qseg;
    int base = 0;
    if (lane == leader) base = atomicAdd(qcount(S, c, x), __popcll(mask));
    base = __shfl(base, leader);
    if (want) {
        const uint64_t below = lane ? mask & ((~0ull) >> (64 - lane)) : 0ull;
        S.qlist[(int64_t)c * S.S + x * S.qseg + base + __popcll(below)] = k;
    }
}

// ---------------------------------------------------------------------
// 2. candidate lists (replaces coretran kNearest, stepper.f90:77-87)
//    outer(f) = { j > f : dist <= r(f) + rtree } by (dist, j)
//    inner(f) = { k < f : dist <= r(k) + rtree } by k
//    fused with each floe's sheltering minima (forcing.f90:162-186; a
//    minimum does not depend on the order its terms come in, and every term
//    is computed from exactly the operands the pair event would use) and the
//    work classification: a floe joins the contact queue if it overlaps a
//    candidate partner or holds contact histories, the bond queue if it owns
//    bond records (or a bond of its own forms this step), the wall queue if
//    it crosses the box.
// ---------------------------------------------------------------------
// Finishing writes of one floe's search (any search shape): list counts,
// per-slot resets, sheltering minima and the per-step resets of its rows
// (fs: the floe's storage slot).
template <bool MIXED>
__device__ __forceinline__ void cand_finish(const DS& S, int k, int fs, const DevParams& P, int no,
                                            int ni, double sha, double shw, bool qc, bool qb,
                                            int lane0, int lanes) {
    const int64_t SS = S.S;
    const uint32_t fl = P.shelter ? EV_SHELTER : 0u;
    for (int s2 = lane0; s2 < no; s2 += lanes) {
        S.evf[s2 * SS + k] = fl;
        if (S.record_forces) {
            S.ssc[s2 * SS + k] = 0.0;
            S.ssb[s2 * SS + k] = 0.0;
        }
    }
    for (int s2 = lane0; s2 < ni; s2 += lanes) S.ibond[s2 * SS + k] = 0;
    if (lane0 == 0) {
        S.ocnt[k] = no;
        S.icnt[k] = ni;
        S.acc[A_SHA * SS + k] = sha;
        S.acc[A_SHW * SS + k] = shw;
        if (!qc) S.hcnt_new[fs] = 0;
        if (!qb) S.bcnt_new[fs] = 0;
        S.wflag[k] = 0;
        S.bgat[k] = 0;
        S.comp[k] = k;
        S.cls[k] = qc ? CLS_CONTACT : 0;   // the fix-up sets CLS_BOND
    }
}

constexpr int SCAN_BATCH = 4;
constexpr int LIST_REGS = 4;

// Candidate superset of one floe by one thread over its (2 SPAN + 1)^2 cell
// block: every floe j within (j > f ? r(f) : r(j)) + rtree + skin.  The cells
// of one row of the block are consecutive in the cell order, so each row is
// one contiguous range of positions; every cell range is read at once
// (indices clamped into the grid, cells outside it masked out afterwards) and
// the rows, taken as one sequence, SCAN_BATCH records at a time.  FILL
// writes the entries at `out`; either way the count is returned.
template <int SPAN, bool FILL>
__device__ __forceinline__ int sup_thread(const DS& S, int k, const GeoRec& me, const DevParams& P,
                                          int64_t* out) {
    constexpr int R = 2 * SPAN + 1;
    const int cx = cell_coord(me.x, P.cell, P.gx), cy = cell_coord(me.y, P.cell, P.gy);
    int2 rg[R][R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int b = min(max(cy - SPAN + t, 0), P.gy - 1);
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int a = min(max(cx - SPAN + u, 0), P.gx - 1);
            rg[t][u] = S.crange[P.cell_base + b * P.gx + a];
        }
    }
    int rlo[R], rlen[R];
    int L = 0;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int b = cy - SPAN + t;
        int lo = INT_MAX, hi = 0;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int a = cx - SPAN + u;
            const bool in = b >= 0 && b < P.gy && a >= 0 && a < P.gx && rg[t][u].y > rg[t][u].x;
            lo = in ? min(lo, rg[t][u].x) : lo;
            hi = in ? max(hi, rg[t][u].y) : hi;
        }
        rlo[t] = lo;
        rlen[t] = lo < hi ? hi - lo : 0;
        L += rlen[t];
    }
    const int f = me.id;
    const double rad = me.r + P.rtree + P.skin, reach = P.rtree + P.skin;
    int n = 0;
    for (int v0 = 0; v0 < L; v0 += SCAN_BATCH) {
        GeoRec g[SCAN_BATCH];
        int qq[SCAN_BATCH];
#pragma unroll
        for (int u = 0; u < SCAN_BATCH; ++u) {
            // position of entry v of the concatenated rows (the last entry
            // stands in past the end; it is not tested twice)
            int rem = min(v0 + u, L - 1), q = 0;
#pragma unroll
            for (int t = 0; t < R; ++t) {
                q = (rem >= 0 && rem < rlen[t]) ? rlo[t] + rem : q;
                rem -= rlen[t];
            }
            qq[u] = q;
            g[u] = S.geo[q];
        }
#pragma unroll
        for (int u = 0; u < SCAN_BATCH; ++u) {
            const int q = qq[u];
            if (v0 + u >= L || q == k) continue;
            const double dx = me.x - g[u].x, dy = me.y - g[u].y;
            const double thr = g[u].id > f ? rad : g[u].r + reach;
            // a superset: the margin only adds entries
            if (dx * dx + dy * dy <= thr * thr * (1.0 + 1e-12)) {
                if (FILL) out[n] = ent(g[u].id, q);
                ++n;
            }
        }
    }
    return n;
}

// (floe, candidate) pairs of one wavefront in LDS: entry, distance (inner:
// -1) and owner lane, then the per-lane pair counts
constexpr int CAND_TCAP = 512;
constexpr int CAND_LDS_PER_WAVE = CAND_TCAP * (8 + 8 + 1) + 64 * 4 + CAND_TCAP * 2;

// Neighbour search of one floe by one thread: its candidate lists are the
// entries of its superset within the query radius, inserted into the
// slot-major rows, then sheltering minima, classification and the per-step
// resets of the floe's rows.  The head of each candidate list is sorted in
// registers while it is built, the rest in its slots in HBM.
//
// WAVE: every lane of the wavefront calls it (act: the lane has a floe).  The
// wavefront's supersets lie one after the other in the pool; they are tested
// together, 64 consecutive entries a round (each lane finds the entry's
// owner among the lanes), and the entries within the query radius are listed
// in LDS (`lds`, CAND_LDS_PER_WAVE bytes for this wavefront) in superset
// order, so each owner's are contiguous.  Each floe inserts its own into its
// lists in that order (as its own scan would), and the sheltering and
// classification terms of all the listed pairs are evaluated together, 64 at
// a time; each floe then takes the minimum of its own terms (a minimum of
// fmin terms does not depend on the order or on which lane computed them).
// Otherwise, or when the pairs outnumber CAND_TCAP, each lane scans its own
// superset SCAN_BATCH records at a time and evaluates its own terms.
// The rows of position k (and its storage slot's counts) that the search
// reads first, loaded at once as soon as k is known.
struct CandPre {
    ShelRec msh;
    int64_t soff;
    int fs, L, hc, bc;
};
__device__ __forceinline__ CandPre cand_pre(const DS& S, int k) {
    CandPre c;
    c.msh = S.shel[k];
    c.fs = S.cval[k];
    c.L = S.scnt[k];
    c.soff = S.soff[k];
    c.hc = S.hcnt_old[c.fs];
    c.bc = S.bcnt_old[c.fs];
    return c;
}

template <bool MIXED, bool WAVE>
__device__ __forceinline__ void cand_thread(const DS& S, int k, const GeoRec& me, const CandPre& pre,
                                            int m, bool act, bool& qc, bool& qb, bool& qw,
                                            unsigned char* lds) {
    using T = typename Prec<MIXED>::T;
    // member m's parameters and step, read at once (m is a valid member
    // index whether or not the lane has a floe); no short-circuit tests
    // between them, which would turn into a chain of single loads
    const DevParams& P = S.params[m];
    const Member Mb = S.members[m];
    const double rtree = P.rtree, pnx = P.nx, pny = P.ny, uw = P.uw, vw = P.vw, blim = P.bond_lim;
    const int coh = P.cohesion, rst = P.restart, shl = P.shelter;
    const uint32_t bthr = P.bond_thr24;
    const uint64_t bseed = P.bond_seed;
    const double mx = act ? me.x : 0.0, my = act ? me.y : 0.0, mr = act ? me.r : 0.0;
    const double rad = mr + rtree;
    const int f = act ? me.id : 0;
    const int first = Mb.first;
    const bool bond_step = act & (Mb.tstep == 1) & (rst != 1) & (coh != 0);
    ShelRec msh;
    msh.hfa = msh.hfw = msh.wx = msh.wy = 0.0;
    const T uwi = T(uw), vwi = T(vw);
    T sha = T(1), shw = T(1);
    int fs = 0;
    int no = 0, ni = 0;
    uint32_t over = 0;
    const int64_t SS = S.S;
    int64_t* oj = S.oj;
    int64_t* ik = S.ik;
    // the first LIST_REGS entries of each list stay in registers while the
    // lists are built (sorted; empty slots hold keys above every entry), the
    // rest are sorted in their slots in HBM
    double rd[LIST_REGS];
    int64_t re[LIST_REGS], ri[LIST_REGS];
#pragma unroll
    for (int t = 0; t < LIST_REGS; ++t) {
        rd[t] = __longlong_as_double(0x7ff0000000000000ll);
        re[t] = INT64_MAX;
        ri[t] = INT64_MAX;
    }
    int L = 0;
    int64_t soff = 0;
    if (act) {
        msh = pre.msh;
        fs = pre.fs;
        qc = pre.hc > 0;
        qb = pre.bc > 0;
        qw = (me.y - me.r < 0) | (me.y + me.r > pny) | (me.x - me.r < 0) | (me.x + me.r > pnx);
        L = pre.L;
        soff = pre.soff;
    }
    // one candidate j at position q and distance d into the floe's lists
    auto insert = [&](int j, int q, double d) {
        if (j > f) {
            if (no >= S.CO) { over |= ERR_OUTER_CAP; return; }
            // insertion by (distance, id) into the register part; the entry
            // that leaves it is the largest so far
            double xd = d;
            int64_t xe = ent(j, q);
#pragma unroll
            for (int t = 0; t < LIST_REGS; ++t) {
                const bool lt = xd < rd[t] || (xd == rd[t] && ent_id(xe) < ent_id(re[t]));
                const double td = rd[t];
                const int64_t te = re[t];
                rd[t] = lt ? xd : td;
                re[t] = lt ? xe : te;
                xd = lt ? td : xd;
                xe = lt ? te : xe;
            }
            if (no >= LIST_REGS) {
                int p = no;
                while (p > LIST_REGS) {
                    double dp = S.od[(p - 1) * SS + k];
                    int64_t ep = oj[(p - 1) * SS + k];
                    if (dp < xd || (dp == xd && ent_id(ep) < ent_id(xe))) break;
                    S.od[p * SS + k] = dp;
                    oj[p * SS + k] = ep;
                    --p;
                }
                S.od[p * SS + k] = xd;
                oj[p * SS + k] = xe;
            }
            ++no;
        } else {
            if (ni >= S.CI) { over |= ERR_INNER_CAP; return; }
            int64_t xe = ent(j, q);
#pragma unroll
            for (int t = 0; t < LIST_REGS; ++t) {
                const bool lt = xe < ri[t];
                const int64_t te = ri[t];
                ri[t] = lt ? xe : te;
                xe = lt ? te : xe;
            }
            if (ni >= LIST_REGS) {
                int p = ni;
                while (p > LIST_REGS) {
                    int64_t ep = ik[(p - 1) * SS + k];
                    if (ep < xe) break;
                    ik[p * SS + k] = ep;
                    --p;
                }
                ik[p * SS + k] = xe;
            }
            ++ni;
        }
    };
    bool flat = false;          // the wavefront's pairs are listed in LDS
    int abase = 0, acount = 0, atotal = 0;
    int64_t* te = reinterpret_cast<int64_t*>(lds);                      // entry, then the water term
    double* td = reinterpret_cast<double*>(lds + 8 * CAND_TCAP);         // distance (inner: -1), then the air term
    uint8_t* to = lds + 16 * CAND_TCAP;                                   // owner lane, then flags | 0x80
    int32_t* tn = reinterpret_cast<int32_t*>(lds + 17 * CAND_TCAP);      // pairs per lane
    // the pairs whose terms the second pass evaluates (indices into the above)
    uint16_t* tx = reinterpret_cast<uint16_t*>(lds + 17 * CAND_TCAP + 64 * 4);
    int nneed = 0;
    if (WAVE) {
        const int lane = (int)(threadIdx.x & 63);
        const uint64_t lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
        int incl = L;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        const int start = incl - L, totalL = __shfl(incl, 63);
        tn[lane] = 0;
        wave_lds_sync();
        bool ovf = false;
        // the floe's switches for the second pass: sheltering on (bit 0), bonds
        // may form this step (bit 1)
        const int pflags = (shl != 0 ? 1 : 0) | (bond_step ? 2 : 0);
        // an entry's owner: the last lane whose superset starts at or before
        // it (lanes without entries share the next one's start); idx < totalL
        auto owner_of = [&](int idx) {
            int own = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1)
                if (__shfl(start, own + step) <= idx) own += step;
            return own;
        };
        auto entry_at = [&](int idx, int own) -> int64_t {
            return S.sup[__shfl(soff, own) + (idx - __shfl(start, own))];
        };
        // each round's superset entries are requested during the round
        // before (a lane past the end reads the last entry again)
        int own_nxt = 0;
        int64_t e_nxt = 0;
        if (totalL > 0) {
            own_nxt = owner_of(min(lane, totalL - 1));
            e_nxt = entry_at(min(lane, totalL - 1), own_nxt);
        }
        for (int c0 = 0; c0 < totalL; c0 += 64) {
            const int idx = c0 + lane;
            const int own = own_nxt;
            const int64_t e = e_nxt;
            if (c0 + 64 < totalL) {
                const int idn = min(idx + 64, totalL - 1);
                own_nxt = owner_of(idn);
                e_nxt = entry_at(idn, own_nxt);
            }
            const double ox = __shfl(mx, own), oy = __shfl(my, own), orad = __shfl(rad, own);
            const double ort = __shfl(rtree, own), orr = __shfl(mr, own);
            const double ohfa = __shfl(msh.hfa, own), ohfw = __shfl(msh.hfw, own);
            const int of = __shfl(f, own), ofl = __shfl(pflags, own);
            bool hit = false, need = false;
            double d = 0.0;
            int j = 0, q = 0;
            if (idx < totalL) {
                q = ent_pos(e);
                const GeoRec o = S.geo[q];
                const double dx = ox - o.x, dy = oy - o.y;
                const double dist2 = dx * dx + dy * dy;
                j = o.id;
                const double thr = j > of ? orad : o.r + ort;
                // far beyond the radius: no square root needed (sqrt is
                // correctly rounded, so it cannot cross a margin of 1e-12)
                if (dist2 <= thr * thr * (1.0 + 1e-12)) {
                    d = sqrt(dist2);
                    hit = d <= thr;
                }
                if (hit) {
                    // The second pass changes nothing for a pair out of
                    // contact, with no bond forming this step, whose two
                    // sheltering terms are exactly 1: the test shelter_min
                    // makes (the same overlap, the floe's own drafts, type T).
                    const double dn = j > of ? orr + o.r - d : o.r + orr - d;
                    const T dnt = T(dn);
                    const T lim = sizeof(T) == 4 ? T(20) : T(40);
                    const bool one_a = dnt < T(0) && T((double)0.18f) * -dnt >= lim * T(ohfa);
                    const bool one_w = dnt < T(0) && T((double)0.18f) * -dnt >= lim * T(ohfw);
                    need = dn > 0 || (ofl & 2) || ((ofl & 1) && !(one_a && one_w));
                }
            }
            const uint64_t mh = __ballot(hit);
            if (atotal + __popcll(mh) > CAND_TCAP) { ovf = true; break; }
            const uint64_t mn = __ballot(need);
            if (hit) {
                const int at = atotal + __popcll(mh & lt);
                te[at] = ent(j, q);
                td[at] = j > of ? d : -1.0;
                to[at] = (uint8_t)own;
                atomicAdd(&tn[own], 1);
                if (need) tx[nneed + __popcll(mn & lt)] = (uint16_t)at;
            }
            atotal += __popcll(mh);
            nneed += __popcll(mn);
        }
        wave_lds_sync();
        if (!ovf) {
            flat = true;
            acount = tn[lane];
            int ai = acount;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int v = __shfl_up(ai, off);
                if (lane >= off) ai += v;
            }
            abase = ai - acount;
            if (act)
                for (int i = abase; i < abase + acount; ++i) {
                    const double dv = td[i];
                    insert(ent_id(te[i]), ent_pos(te[i]), dv);
                }
        }
    }
    if (!flat && act) {
        // (the wavefront search is the common path: this one keeps fewer
        // records in flight, which keeps the kernel's registers down)
        constexpr int B = WAVE ? 2 : SCAN_BATCH;
        const int64_t* se = S.sup + soff;
        for (int v0 = 0; v0 < L; v0 += B) {
            GeoRec g[B];
            int qq[B];
#pragma unroll
            for (int u = 0; u < B; ++u)
                qq[u] = ent_pos(se[min(v0 + u, L - 1)]);   // the last entry stands in past the end
#pragma unroll
            for (int u = 0; u < B; ++u) g[u] = S.geo[qq[u]];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if (v0 + u >= L) continue;
                const GeoRec* o = &g[u];
                const double dx = me.x - o->x, dy = me.y - o->y;
                const double dist2 = dx * dx + dy * dy;
                const double thr = o->id > f ? rad : o->r + rtree;
                if (dist2 > thr * thr * (1.0 + 1e-12)) continue;
                const double d = sqrt(dist2);
                if (!(d <= thr)) continue;
                insert(o->id, qq[u], d);
            }
        }
    }
    // sheltering and classification over the pairs found.  A minimum and an
    // "any" do not depend on the order of their terms.
    if (flat) {
        const int lane = (int)(threadIdx.x & 63);
        wave_lds_sync();
        for (int c0 = 0; c0 < nneed; c0 += 64) {
            const bool valid = c0 + lane < nneed;
            const int idx = valid ? (int)tx[c0 + lane] : 0;
            int64_t e = 0;
            double d = -1.0;
            int own = lane;
            if (valid) { e = te[idx]; d = td[idx]; own = to[idx]; }
            // the pair's floe's member parameters, from its lane
            const int ko = __shfl(k, own);
            const int first_o = __shfl(first, own);
            const int bst_o = __shfl((int)bond_step, own), shl_o = __shfl(shl, own);
            const double blim_o = __shfl(blim, own), uw_o = __shfl(uw, own), vw_o = __shfl(vw, own);
            const uint32_t thr_o = (uint32_t)__shfl((int)bthr, own);
            const uint64_t seed_o = __shfl(bseed, own);
            T ta = T(1), tw = T(1);
            uint32_t fl = 0;
            if (valid) {
                const int q = ent_pos(e);
                const GeoRec mo = S.geo[ko];
                const GeoRec o = S.geo[q];
                const ShelRec oh = S.shel[q];
                const ShelRec mh = S.shel[ko];
                const bool outer = d >= 0.0;
                double dd = d;
                if (!outer) {
                    const double dx = mo.x - o.x, dy = mo.y - o.y;
                    dd = sqrt(dx * dx + dy * dy);
                }
                // outer: pair (f, j), f the outer floe i of the event; inner:
                // pair (j, f), f the inner floe j
                const double deltan = outer ? mo.r + o.r - dd : o.r + mo.r - dd;
                if (deltan > 0) fl |= 1u;
                if (outer && bst_o && deltan >= -blim_o &&
                    bond_keep(mo.id - first_o, ent_id(e) - first_o, seed_o, thr_o))
                    fl |= 2u;
                if (shl_o) {
                    // outer: floe f takes hsf*(j,f), evaluated with -cosa,
                    // -sina of the direction f -> j; inner: cosa/sina point
                    // from j to f and f takes hsf*(j,f) with j's wind
                    const double nx = outer ? o.x - mo.x : mo.x - o.x;
                    const double ny = outer ? o.y - mo.y : mo.y - o.y;
                    const double cosa = nx / dd, sina = ny / dd;
                    const T dn = T(deltan), c0t = T(cosa), s0t = T(sina);
                    const T c = outer ? -c0t : c0t, sn = outer ? -s0t : s0t;
                    const T ux = outer ? T(mh.wx) : T(oh.wx), uy = outer ? T(mh.wy) : T(oh.wy);
                    ta = shelter_min(T(1), T(oh.hfa), T(mh.hfa), dn, c, sn, ux, uy);
                    tw = shelter_min(T(1), T(oh.hfw), T(mh.hfw), dn, c, sn, T(uw_o), T(vw_o));
                }
            }
            if (valid) {
                td[idx] = (double)ta;
                te[idx] = __double_as_longlong((double)tw);
                to[idx] = (uint8_t)(fl | 0x80u);
            }
        }
        wave_lds_sync();
        if (act) {
            for (int i = abase; i < abase + acount; ++i) {
                if (!(to[i] & 0x80u)) continue;      // not evaluated: terms 1, no flags
                sha = dmin(sha, T(td[i]));
                shw = dmin(shw, T(__longlong_as_double(te[i])));
                qc = qc || (to[i] & 1u);
                qb = qb || (to[i] & 2u);
            }
        }
    } else if (act) {
        auto outer_terms = [&](int64_t e, double d) {
            const int j = ent_id(e), q = ent_pos(e);
            const GeoRec o = S.geo[q];
            // pair (f, j): f is the outer floe i of the event
            const double deltan = me.r + o.r - d;
            if (deltan > 0) qc = true;
            if (!qb && bond_step && deltan >= -blim &&
                bond_keep(f - first, j - first, bseed, bthr))
                qb = true;
            if (shl) {
                // floe f takes hsf*(j,f), evaluated with -cosa, -sina
                const double cosa = (o.x - me.x) / d, sina = (o.y - me.y) / d;
                const ShelRec oh = S.shel[q];
                const T dn = T(deltan), c = T(cosa), sn = T(sina);
                sha = shelter_min(sha, T(oh.hfa), T(msh.hfa), dn, -c, -sn, T(msh.wx), T(msh.wy));
                shw = shelter_min(shw, T(oh.hfw), T(msh.hfw), dn, -c, -sn, uwi, vwi);
            }
        };
        auto inner_terms = [&](int64_t e) {
            const int q = ent_pos(e);
            const GeoRec o = S.geo[q];
            const double dx = me.x - o.x, dy = me.y - o.y;
            const double d = sqrt(dx * dx + dy * dy);
            // pair (j, f): f is the inner floe j of the event
            const double deltan = o.r + me.r - d;
            if (deltan > 0) qc = true;
            if (shl) {
                // cosa/sina point from j to f; floe f takes hsf*(j,f) with j's wind
                const double cosa = (me.x - o.x) / d, sina = (me.y - o.y) / d;
                const ShelRec oh = S.shel[q];
                const T dn = T(deltan), c = T(cosa), sn = T(sina);
                sha = shelter_min(sha, T(oh.hfa), T(msh.hfa), dn, c, sn, T(oh.wx), T(oh.wy));
                shw = shelter_min(shw, T(oh.hfw), T(msh.hfw), dn, c, sn, uwi, vwi);
            }
        };
#pragma unroll
        for (int s2 = 0; s2 < LIST_REGS; ++s2)
            if (s2 < no) outer_terms(re[s2], rd[s2]);
        for (int s2 = LIST_REGS; s2 < no; ++s2) outer_terms(oj[s2 * SS + k], S.od[s2 * SS + k]);
#pragma unroll
        for (int s2 = 0; s2 < LIST_REGS; ++s2)
            if (s2 < ni) inner_terms(ri[s2]);
        for (int s2 = LIST_REGS; s2 < ni; ++s2) inner_terms(ik[s2 * SS + k]);
    }
    if (!act) return;
    // the register part of the lists goes to its slots
#pragma unroll
    for (int s2 = 0; s2 < LIST_REGS; ++s2) {
        if (s2 < no) {
            oj[s2 * SS + k] = re[s2];   // (distances are only read back beyond the register part)
        }
        if (s2 < ni) ik[s2 * SS + k] = ri[s2];
    }
    if (over) set_err(S, over);
    cand_finish<MIXED>(S, k, fs, P, no, ni, double(sha), double(shw), qc, qb, 0, 1);
}

// Superset build, thread per floe (members that are not dense; span 1 or
// 2 cells on each side): counts (FILL false) or entries and the build
// positions x0 (FILL true).  The count pass lists the dense members' floes in
// queue 3 for k_sup_dense and k_cand_dense.  Only on build steps.
template <bool FILL>
__global__ __launch_bounds__(FLOE_BLOCK) void k_sup(DS S) {
    if (!*S.rebuild || *S.err) return;
    const int k = xcd_block() * blockDim.x + threadIdx.x;
    bool qd = false;
    if (k < S.N) {
        const GeoRec me = S.geo[k];
        if (me.member < 0) {
            if (!FILL) S.scnt[k] = 0;
        } else {
            const DevParams& P = S.params[me.member];
            if (P.dense) {
                qd = !FILL;
            } else if (FILL) {
                if (P.span == 2) sup_thread<2, true>(S, k, me, P, S.sup + S.soff[k]);
                else sup_thread<1, true>(S, k, me, P, S.sup + S.soff[k]);
                S.x0[k] = make_double2(me.x, me.y);
            } else {
                S.scnt[k] = P.span == 2 ? sup_thread<2, false>(S, k, me, P, nullptr)
                                        : sup_thread<1, false>(S, k, me, P, nullptr);
            }
        }
    }
    enqueue(S, qd, 3, k);
}

// Superset offsets: per-tile totals, their exclusive scan (one workgroup;
// the total must fit the superset pool), then offsets within each tile.
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* tmp, int64_t& total) {
    const int t = (int)threadIdx.x, n = (int)blockDim.x;
    tmp[t] = v;
    __syncthreads();
    for (int off = 1; off < n; off <<= 1) {
        const int64_t add = t >= off ? tmp[t - off] : 0;
        __syncthreads();
        tmp[t] += add;
        __syncthreads();
    }
    total = tmp[n - 1];
    const int64_t incl = tmp[t];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(SCAN_TILE) void k_scan_tiles(DS S) {
    __shared__ int64_t tmp[SCAN_TILE];
    if (!*S.rebuild || *S.err) return;
    const int k = blockIdx.x * SCAN_TILE + threadIdx.x;
    int64_t tot;
    block_exclusive_scan(k < S.N ? S.scnt[k] : 0, tmp, tot);
    if (threadIdx.x == 0) S.sblk[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_TILE) void k_scan_total(DS S, int ntiles) {
    __shared__ int64_t tmp[SCAN_TILE];
    if (!*S.rebuild || *S.err) return;
    int64_t base = 0;
    for (int t0 = 0; t0 < ntiles; t0 += SCAN_TILE) {
        const int t = t0 + (int)threadIdx.x;
        const int64_t v = t < ntiles ? S.sblk[t] : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan(v, tmp, tot);
        if (t < ntiles) S.sblk[t] = base + ex;
        base += tot;
    }
    if (threadIdx.x == 0) {
        S.sblk[ntiles] = base;
        if (base > S.sup_cap) set_err(S, ERR_SUP_CAP);
    }
}

__global__ __launch_bounds__(SCAN_TILE) void k_scan_offsets(DS S) {
    __shared__ int64_t tmp[SCAN_TILE];
    if (!*S.rebuild || *S.err) return;
    const int k = blockIdx.x * SCAN_TILE + threadIdx.x;
    int64_t tot;
    const int64_t ex = block_exclusive_scan(k < S.N ? S.scnt[k] : 0, tmp, tot);
    if (k < S.N) S.soff[k] = S.sblk[blockIdx.x] + ex;
}

template <bool MIXED>
__global__ __launch_bounds__(FLOE_BLOCK) void k_cand(DS S) {
    const int k = xcd_block() * blockDim.x + threadIdx.x;
    bool qc = false, qb = false, qw = false;
    // every row the search starts from is requested before anything is
    // tested (reads only: harmless for free slots and failed steps)
    const bool inr = k < S.N;
    GeoRec me;
    CandPre pre;
    if (inr) {
        me = S.geo[k];
        pre = cand_pre(S, k);
    }
    const bool slot_used = inr && me.member >= 0;
    const int m = slot_used ? me.member : 0;
    const bool dense = S.params[m].dense != 0;
    const bool ok = !*S.err;
    // floes of dense members: k_cand_dense
    extern __shared__ unsigned char cand_lds[];
    const bool act = slot_used & ok & !dense;
    cand_thread<MIXED, true>(S, k, me, pre, m, act, qc, qb, qw,
                             cand_lds + (threadIdx.x >> 6) * CAND_LDS_PER_WAVE);
    if (inr && ok && !slot_used) {
        S.ocnt[k] = 0;
        S.icnt[k] = 0;
    }
    enqueue(S, qc, 0, k);
    enqueue(S, qb, 1, k);
    enqueue(S, qw, 2, k);
}

// Dense members (many floes per cell): a thread per floe, over its superset
// as the build ordered it (outer entries by distance at the build, then inner
// ones by id).  Inner candidates therefore come in list order and are
// appended; outer ones come in nearly the order of their current distance
// (floes move little between builds), so each is inserted among the last
// DENSE_WIN entries, which stay in registers (most simply take their top),
// and only the rare entry that belongs below them is inserted into the part
// of the list in HBM.
// The lists are slot-major (slot s of floe k at s * S + k), so a thread that
// stored each entry as it found it would touch a line per entry, 64 lines per
// store of a wavefront.  Instead the scan takes the outer part of the
// supersets in one pass and the inner part in another (one after the other,
// or in two wavefronts over the same floes), so that the lanes' lists grow
// nearly in step, and each thread stages its rows in a ring of its own in
// LDS; the wavefront writes a row out for all its floes at once when every
// thread still scanning has passed it, or when a thread needs the room.
// A row a thread reaches after its wavefront wrote it out goes straight to
// HBM; a thread that inserts below its register part first writes its staged
// outer rows out and keeps its outer list in HBM from then on.
// Sheltering terms are needed only for pairs close enough that
// exp(-0.18 |dn/cos| / hf) still counts (shelter_min); the threads list those
// pairs in LDS (slots taken by ballot, no atomic) and the wavefront evaluates
// them 64 at a time, each floe taking the minimum of its own (an LDS atomic:
// a minimum does not depend on the order of its terms).  A superset left in
// scan order (longer than SUP_ORDER_CAP) goes to the thread search
// (k_cand_dense_over).
constexpr int DENSE_OVER = -1;   // ocnt of a dense floe left to the thread search
constexpr int DENSE_WIN = 4;     // outer candidates in registers
constexpr int DENSE_NEAR = 512;  // near pairs listed per wavefront
constexpr int DENSE_BATCH = 5;   // superset entries a thread loads at once (mixed precision)
constexpr int DENSE_RING = 16;   // list rows a thread stages in LDS
constexpr int DENSE_WAVES_PER_SIMD = 3;   // the dense search's occupancy bound
static_assert((DENSE_RING & (DENSE_RING - 1)) == 0 && DENSE_RING >= DENSE_BATCH + DENSE_WIN,
              "a ring of whole rows with room for a batch past the register part");

__device__ __forceinline__ void bitonic_inner(int64_t* ke, int n, int lane) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = lane; i < n; i += 64) {
                const int pj = i ^ stride;
                if (pj > i) {
                    const bool up = (i & size) == 0;
                    const int64_t e1 = ke[i], e2 = ke[pj];
                    if ((e1 > e2) == up) { ke[i] = e2; ke[pj] = e1; }
                }
            }
            __syncthreads();
        }
    }
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dmin(v, (T)__shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ int pow2_at_least(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Superset build of the dense members' floes (queue 3), one wavefront per
// floe over its (2 span + 1)^2 cell block, 64 entries at a time: counts, or
// the entries and the build positions.  The filled superset is then ordered
// as the dense search reads it: outer entries (partner id above the floe's)
// first, by (distance at the build, id), then inner entries by id; each
// entry goes to its rank, counted over the floe's entries in LDS.  A
// superset longer than SUP_ORDER_CAP stays in scan order (sout = -1).  Only
// on build steps.
constexpr int SUP_ORDER_CAP = 256;   // (16 bytes of LDS per entry: 4 KB per wavefront)
template <bool FILL>
__global__ __launch_bounds__(64) void k_sup_dense(DS S) {
    __shared__ double od_[FILL ? SUP_ORDER_CAP : 1];
    __shared__ int64_t oe_[FILL ? SUP_ORDER_CAP : 1];
    if (!*S.rebuild || *S.err) return;
    const int lane = threadIdx.x;
    const QueueView qv = queue_view(S, 3);
    const int nq = qv.total;
    const int nb = (int)(gridDim.x / XCD_COUNT);
    const int xcd = (int)(blockIdx.x % XCD_COUNT), bi = (int)(blockIdx.x / XCD_COUNT);
    const int per = (nq + XCD_COUNT - 1) / XCD_COUNT;
    const int t1 = min(nq, xcd * per + per);
    const uint64_t lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    for (int t = xcd * per + bi; t < t1; t += nb) {
        const int k = queue_at(S, 3, qv, t);
        const GeoRec me = S.geo[k];
        const DevParams& P = S.params[me.member];
        const double rad = me.r + P.rtree + P.skin, reach = P.rtree + P.skin;
        const int f = me.id;
        const int sp = P.span;
        const int cx = cell_coord(me.x, P.cell, P.gx), cy = cell_coord(me.y, P.cell, P.gy);
        int64_t* out = FILL ? S.sup + S.soff[k] : nullptr;
        int n = 0;
        const int a0 = max(cx - sp, 0), a1 = min(cx + sp, P.gx - 1);
        for (int b = cy - sp; b <= cy + sp; ++b) {
            if (b < 0 || b >= P.gy) continue;
            // the cells of one row are consecutive in the cell order
            int lo = INT_MAX, hi = 0;
            for (int a = a0; a <= a1; ++a) {
                const int2 rg = S.crange[P.cell_base + b * P.gx + a];
                if (rg.y > rg.x) { lo = min(lo, rg.x); hi = max(hi, rg.y); }
            }
            for (int base = lo; base < hi; base += 64) {
                const int q = base + lane;
                bool in = false;
                int oid = 0;
                if (q < hi && q != k) {
                    const GeoRec o = S.geo[q];
                    const double dx = me.x - o.x, dy = me.y - o.y;
                    oid = o.id;
                    const double thr = o.id > f ? rad : o.r + reach;
                    in = dx * dx + dy * dy <= thr * thr * (1.0 + 1e-12);
                }
                const uint64_t mi = __ballot(in);
                if (FILL && in) out[n + (int)__popcll(mi & lt)] = ent(oid, q);
                n += (int)__popcll(mi);
            }
        }
        if (FILL) {
            if (n <= SUP_ORDER_CAP) {
                // outer entries carry their build distance, inner ones -1
                int nout = 0;
                for (int v = lane; v < n; v += 64) {
                    const int64_t e = out[v];
                    const GeoRec o = S.geo[ent_pos(e)];
                    const double dx = me.x - o.x, dy = me.y - o.y;
                    const bool outer = ent_id(e) > f;
                    od_[v] = outer ? sqrt(dx * dx + dy * dy) : -1.0;
                    oe_[v] = e;
                    nout += (int)__popcll(__ballot(outer));
                }
                nout = __shfl(nout, 0);
                wave_lds_sync();
                for (int v = lane; v < n; v += 64) {
                    const double dv = od_[v];
                    const int64_t ev = oe_[v];
                    int rank = 0;
                    if (dv >= 0.0) {
                        for (int u = 0; u < n; ++u) {
                            const double du = od_[u];
                            rank += (du >= 0.0 && (du < dv || (du == dv && oe_[u] < ev))) ? 1 : 0;
                        }
                    } else {
                        rank = nout;
                        for (int u = 0; u < n; ++u) rank += (od_[u] < 0.0 && oe_[u] < ev) ? 1 : 0;
                    }
                    out[rank] = ev;
                }
                if (lane == 0) S.sout[k] = nout;
                wave_lds_sync();
            } else if (lane == 0) {
                S.sout[k] = -1;
            }
        }
        if (lane == 0) {
            if (FILL) S.x0[k] = make_double2(me.x, me.y);
            else S.scnt[k] = n;
        }
    }
}

// A pair whose sheltering terms are needed (its distance is recomputed from
// the two positions, the same expression as the scan's).
struct DenseNear {
    int32_t q;                   // partner position
    uint8_t lane, outer;
    uint16_t pad_;
};

// Order-preserving map of doubles to unsigned keys (a < b iff key(a) < key(b)
// for non-NaN values), for minima taken with integer atomics.
__device__ __forceinline__ uint64_t ord_key(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}
__device__ __forceinline__ double ord_val(uint64_t k) {
    return __longlong_as_double((long long)((k >> 63) ? (k & ~(1ull << 63)) : ~k));
}

// W2: two wavefronts over the same 64 floes, one per half of the
// supersets (for dense queues that fit on the GPU at once); else one
// wavefront takes both halves, one after the other.  (DENSE_WAVES_PER_SIMD
// wavefronts per SIMD: the registers of a batch of partner rows and of two
// sets of superset entries, 5 in mixed precision and 4 in FP64, and the LDS
// of the rings and the near-pair list fit them without spilling.)
template <bool MIXED, bool W2>
__global__ __launch_bounds__(W2 ? 128 : 64, DENSE_WAVES_PER_SIMD) void k_cand_dense(DS S) {
    using T = typename Prec<MIXED>::T;
    __shared__ DenseNear nr[DENSE_NEAR];
    __shared__ uint64_t mina[64], minw[64];
    // staged rows of each wavefront: row r of lane l at (r % DENSE_RING) * 64 + l
    __shared__ int64_t ring2_[W2 ? 2 : 1][DENSE_RING * 64];
    __shared__ int xni[W2 ? 64 : 1];    // the inner wavefront's counts, flags and errors
    __shared__ uint32_t xfl[W2 ? 64 : 1], xover[W2 ? 64 : 1];
    __shared__ int ncnt_[2], err0;      // near pairs listed by each wavefront
    // (the error word is read once for the block: a block that sets it may
    // finish between the two wavefronts' starts, and both must take the
    // barriers below or neither)
    if (threadIdx.x == 0) err0 = *S.err;
    __syncthreads();
    if (err0) return;
    // (W2: wavefront 0 the outer lists, 1 the inner ones)
    const int lane = threadIdx.x & 63, wv = W2 ? (int)(threadIdx.x >> 6) : 0;
    int64_t* const ring_ = ring2_[wv];
    int k = -1;
    {
        const QueueView qv = queue_view(S, 3);
        const int per = (qv.total + XCD_COUNT - 1) / XCD_COUNT;
        const int local = (int)(blockIdx.x / XCD_COUNT) * 64 + lane;
        const int t = (int)(blockIdx.x % XCD_COUNT) * per + local;
        if (local < per && t < qv.total) k = queue_at(S, 3, qv, t);
    }
    GeoRec me;
    me.x = me.y = me.r = 0.0;
    me.id = 0;
    me.member = -1;
    int L = 0, nout = 0;
    int64_t soff = 0;
    if (k >= 0) {
        me = S.geo[k];
        if (me.member >= 0) {   // (-1: member past its stop step, nothing to do)
            L = S.scnt[k];
            nout = S.sout[k];
            soff = S.soff[k];
        }
    }
    bool act = k >= 0 && me.member >= 0;
    if (act && nout < 0) {
        if (wv == 0) S.ocnt[k] = DENSE_OVER;   // superset in scan order: the thread search
        act = false;
    }
    const uint64_t one_key = ord_key(1.0);
    if (wv == 0) {
        mina[lane] = minw[lane] = one_key;
    }
    const int m = act ? me.member : 0;
    const DevParams& P = S.params[m];
    // the parameters every entry of the scan reads, held in registers (the
    // list stores below could otherwise alias them); the bond parameters, the
    // wind and the floe's sheltering record are read again where needed, so
    // that the registers go to the batch of partners
    const double rtree = P.rtree;
    const bool shelter = P.shelter != 0;
    const double rad = me.r + rtree;
    const int f = me.id;
    const int first = S.members[m].first;
    const bool bond_step = act && S.members[m].tstep == 1 && P.restart != 1 && P.cohesion;
    const T lim = sizeof(T) == 4 ? T(20) : T(40);
    T lha = T(0), lhw = T(0);   // lim times the floe's two drafts
    int fs = 0;
    bool qc = false, qb = false;
    if (act) {
        const ShelRec sh = S.shel[k];
        lha = lim * T(sh.hfa);
        lhw = lim * T(sh.hfw);
        fs = S.cval[k];
        qc = S.hcnt_old[fs] > 0;
        qb = S.bcnt_old[fs] > 0;
    }
    if (!act) L = nout = 0;
    T sha = T(1), shw = T(1);
    __syncthreads();
    const int64_t SS = S.S;
    int64_t* oj = S.oj;
    const int64_t* se = S.sup + soff;
    // the last min(no, DENSE_WIN) outer entries, ascending: slots [no - nw, no)
    double wd[DENSE_WIN];
    int64_t we[DENSE_WIN];
#pragma unroll
    for (int w = 0; w < DENSE_WIN; ++w) {
        wd[w] = __longlong_as_double(0x7ff0000000000000ll);
        we[w] = INT64_MAX;
    }
    int no = 0, ni = 0, nw = 0;
    // rows below bo (outer) and bi (inner) are in HBM, a lane's rows from
    // there on in its ring; od: this lane's outer list is all in HBM
    int bo = 0, bi = 0;
    bool od = false;
    uint32_t over = 0;
    // near pairs: this wavefront's part of the list (W2: half each) and the
    // count of pairs listed so far
    const int ncap = W2 ? DENSE_NEAR / 2 : DENSE_NEAR, nofs = wv * (DENSE_NEAR / 2);
    int nbase = 0;
    const uint64_t lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    auto rslot = [&](int r) { return (r & (DENSE_RING - 1)) * 64 + lane; };
    auto less = [](double ad, int64_t ae, double bd, int64_t be) {
        return ad < bd || (ad == bd && ent_id(ae) < ent_id(be));
    };
    // (the part of the list in HBM keeps no distances: the rare entry that
    // belongs below the register part compares with a slot's entry by its
    // distance recomputed from the positions, the same expression)
    auto dist_of = [&](int64_t e) {
        const GeoRec o = S.geo[ent_pos(e)];
        const double dx = me.x - o.x, dy = me.y - o.y;
        return sqrt(dx * dx + dy * dy);
    };
    auto put_outer = [&](int s, int64_t e) {
        if (od || s < bo) oj[(int64_t)s * SS + k] = e;
        else ring_[rslot(s)] = e;
    };
    auto insert_outer = [&](double xd, int64_t xe) {
        if (no >= S.CO) { over |= ERR_OUTER_CAP; return; }
        if (nw == DENSE_WIN && less(xd, xe, wd[0], we[0])) {
            if (!od) {
                // the staged rows [bo, no - DENSE_WIN) out first
                for (int s = bo; s < no - DENSE_WIN; ++s) oj[(int64_t)s * SS + k] = ring_[rslot(s)];
                od = true;
            }
            // below the register part: into the part in HBM, whose slots
            // [0, no - DENSE_WIN) move up by one from x's place on
            int p = no - DENSE_WIN;
            while (p > 0) {
                const int64_t ep = oj[(int64_t)(p - 1) * SS + k];
                if (!less(xd, xe, dist_of(ep), ep)) break;
                oj[(int64_t)p * SS + k] = ep;
                --p;
            }
            oj[(int64_t)p * SS + k] = xe;
            ++no;
            return;
        }
        if (nw == DENSE_WIN) {
            // the smallest register entry is final: to its slot
            put_outer(no - DENSE_WIN, we[0]);
            // (most entries come above the whole register part: they take
            // its top without the insertion below)
            const bool top = !less(xd, xe, wd[DENSE_WIN - 1], we[DENSE_WIN - 1]);
#pragma unroll
            for (int w = 0; w < DENSE_WIN - 1; ++w) { wd[w] = wd[w + 1]; we[w] = we[w + 1]; }
            if (top) {
                wd[DENSE_WIN - 1] = xd;
                we[DENSE_WIN - 1] = xe;
                ++no;
                return;
            }
            wd[DENSE_WIN - 1] = __longlong_as_double(0x7ff0000000000000ll);
            we[DENSE_WIN - 1] = INT64_MAX;
            --nw;
        }
        // insertion into the sorted register part (empty slots hold keys
        // above every entry)
#pragma unroll
        for (int w = 0; w < DENSE_WIN; ++w) {
            const bool lt = less(xd, xe, wd[w], we[w]);
            const double td = wd[w];
            const int64_t te = we[w];
            wd[w] = lt ? xd : td;
            we[w] = lt ? xe : te;
            xd = lt ? td : xd;
            xe = lt ? te : xe;
        }
        ++nw;
        ++no;
    };
    // row bo (bi) out to HBM for every lane that has it staged (the whole
    // wavefront takes part)
    auto out_outer = [&]() {
        if (!od && no - nw > bo) oj[(int64_t)bo * SS + k] = ring_[rslot(bo)];
        ++bo;
    };
    auto out_inner = [&]() {
        if (ni > bi) S.ik[(int64_t)bi * SS + k] = ring_[rslot(bi)];
        ++bi;
    };
    constexpr int B = MIXED ? DENSE_BATCH : DENSE_BATCH - 1;
    // one pass over a part of every lane's superset: outer entries [0, nout),
    // then inner ones [nout, L), entry t of the part at the same step
    auto scan = [&](auto outer_part) {
        constexpr bool outer = decltype(outer_part)::value;
        const int vb = outer ? 0 : nout;
        const int n = outer ? nout : L - nout;
        int nmax = n;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) nmax = max(nmax, __shfl_xor(nmax, off));
        // a batch's superset entries are requested with the batch before, in
        // the other of two register sets (a copy from one set to the other
        // would wait for the loads); the last entry stands in past the end
        int64_t ea[B], eb[B];
#pragma unroll
        for (int u = 0; u < B; ++u) ea[u] = eb[u] = 0;
        if (n > 0) {
#pragma unroll
            for (int u = 0; u < B; ++u) ea[u] = se[vb + min(u, n - 1)];
        }
        auto batch = [&](const int t0, const int64_t (&ev)[B], int64_t (&en)[B]) {
            const bool run = t0 < n;
            // rows out: each row that every lane still scanning has passed,
            // and those that make room for this batch (at most B rows, the
            // register part's included)
            for (;;) {
                const int c = outer ? no - nw : ni;
                const int b = outer ? bo : bi;
                const bool mine = outer ? !od : true;
                if (!__ballot(mine && c > b)) break;
                if (!__ballot(run && mine && (outer ? no : ni) + B > b + DENSE_RING) &&
                    __ballot(run && mine && c <= b))
                    break;
                if (outer) out_outer(); else out_inner();
            }
            // (every lane takes part in the batch, so that the near-pair
            // count below stays the same in all of them)
            GeoRec g[B];
            if (run) {
#pragma unroll
                for (int u = 0; u < B; ++u) g[u] = S.geo[ent_pos(ev[u])];
                if (t0 + B < n) {
#pragma unroll
                    for (int u = 0; u < B; ++u) en[u] = se[vb + min(t0 + B + u, n - 1)];
                }
            }
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const GeoRec& o = g[u];
                const double dx = me.x - o.x, dy = me.y - o.y;
                const double dist2 = dx * dx + dy * dy;
                const double thr = outer ? rad : o.r + rtree;
                // far beyond the radius: no square root needed
                bool inr = run && t0 + u < n && !(dist2 > thr * thr * (1.0 + 1e-12));
                double d = 0.0;
                if (inr) d = sqrt(dist2);
                inr = inr && d <= thr;
                // outer: pair (f, j) with f the outer floe i of the event;
                // inner: pair (j, f) with f the inner floe j
                const double deltan = outer ? me.r + o.r - d : o.r + me.r - d;
                const T dn = T(deltan);
                bool near = false;
                if (inr) {
                    if (deltan > 0) qc = true;
                    if (outer && bond_step && deltan >= -P.bond_lim &&
                        bond_keep(f - first, ent_id(ev[u]) - first, P.bond_seed, P.bond_thr24))
                        qb = true;
                    near = shelter && (!(dn < T(0) && T((double)0.18f) * -dn >= lha) ||
                                       !(dn < T(0) && T((double)0.18f) * -dn >= lhw));
                }
                // near pairs take consecutive slots of the wavefront's part of
                // the list, counted by ballot (the same count in every lane)
                const uint64_t mn = __ballot(near);
                const int slot = nbase + (int)__popcll(mn & lt);
                nbase += (int)__popcll(mn);
                if (near) {
                    {
                        if (slot < ncap) {
                            DenseNear r;
                            r.q = ent_pos(ev[u]); r.lane = (uint8_t)lane; r.outer = outer ? 1 : 0;
                            r.pad_ = 0;
                            nr[nofs + slot] = r;
                        } else {
                            // list full: this thread takes the terms itself
                            const double nx = outer ? o.x - me.x : me.x - o.x;
                            const double ny = outer ? o.y - me.y : me.y - o.y;
                            const double cosa = nx / d, sina = ny / d;
                            const ShelRec ms = S.shel[k];
                            const ShelRec oh = S.shel[ent_pos(ev[u])];
                            const T c0t = T(cosa), s0t = T(sina);
                            const T c = outer ? -c0t : c0t, sn = outer ? -s0t : s0t;
                            const T ux = outer ? T(ms.wx) : T(oh.wx), uy = outer ? T(ms.wy) : T(oh.wy);
                            sha = shelter_min(sha, T(oh.hfa), T(ms.hfa), dn, c, sn, ux, uy);
                            shw = shelter_min(shw, T(oh.hfw), T(ms.hfw), dn, c, sn, T(P.uw), T(P.vw));
                        }
                    }
                }
                if (!inr) continue;
                if (outer) {
                    insert_outer(d, ev[u]);
                } else if (ni < S.CI) {
                    if (ni < bi) S.ik[(int64_t)ni * SS + k] = ev[u];
                    else ring_[rslot(ni)] = ev[u];
                    ++ni;
                } else {
                    over |= ERR_INNER_CAP;
                }
            }
            if (outer && run && t0 + B >= n) {
                // the lane's last outer entries: its register part is final
#pragma unroll
                for (int w = 0; w < DENSE_WIN; ++w)
                    if (w < nw) put_outer(no - nw + w, we[w]);
                nw = 0;
            }
        };
        for (int t0 = 0; t0 < nmax; t0 += 2 * B) {
            batch(t0, ea, eb);
            if (t0 + B < nmax) batch(t0 + B, eb, ea);
        }
    };
    if (wv == 0) {
        scan(std::true_type{});
        // every staged outer row out (each lane's register part went to its
        // rows with its last batch): the ring is free for the inner half
        while (__ballot(!od && no > bo)) out_outer();
    }
    if (!W2 || wv == 1) {
        scan(std::false_type{});
        while (__ballot(ni > bi)) out_inner();
        if (W2) {
            xni[lane] = ni;
            xfl[lane] = qc ? 1u : 0u;
            xover[lane] = over;
        }
    }
    if (lane == 0) ncnt_[wv] = min(nbase, ncap);
    // minima a thread took itself (near-pair list full) join the LDS ones
    if (sha < T(1)) atomicMin((unsigned long long*)&mina[lane], (unsigned long long)ord_key((double)sha));
    if (shw < T(1)) atomicMin((unsigned long long*)&minw[lane], (unsigned long long)ord_key((double)shw));
    // sheltering terms of the listed near pairs, across the lanes (every lane
    // takes part in each round; the listing thread's floe comes by shuffle)
    ShelRec msh;
    msh.hfa = msh.hfw = msh.wx = msh.wy = 0.0;
    if (act) msh = S.shel[k];
    __syncthreads();
    // (W2: the two wavefronts take alternate rounds of 64 pairs)
    const int nn0 = ncnt_[0], nn = nn0 + (W2 ? ncnt_[1] : 0);
    for (int i0 = wv * 64; i0 < nn; i0 += W2 ? 128 : 64) {
        const int i = i0 + lane;
        DenseNear r;
        r.q = 0; r.lane = 0; r.outer = 0; r.pad_ = 0;
        if (i < nn) r = nr[i < nn0 ? i : DENSE_NEAR / 2 + (i - nn0)];
        const int ow = r.lane;
        const double ox = __shfl(me.x, ow), oy = __shfl(me.y, ow), orr = __shfl(me.r, ow);
        const double ohfa = __shfl(msh.hfa, ow), ohfw = __shfl(msh.hfw, ow);
        const double owx = __shfl(msh.wx, ow), owy = __shfl(msh.wy, ow);
        const double ouw = __shfl(P.uw, ow), ovw = __shfl(P.vw, ow);
        if (i < nn) {
            const GeoRec o = S.geo[r.q];
            const ShelRec oh = S.shel[r.q];
            const bool outer = r.outer != 0;
            const double ddx = ox - o.x, ddy = oy - o.y;
            const double d = sqrt(ddx * ddx + ddy * ddy);
            // outer: f takes hsf*(j,f) with -cosa, -sina of f -> j and its own
            // wind; inner: cosa/sina from j to f and j's wind
            const double deltan = outer ? orr + o.r - d : o.r + orr - d;
            const double nx = outer ? o.x - ox : ox - o.x;
            const double ny = outer ? o.y - oy : oy - o.y;
            const double cosa = nx / d, sina = ny / d;
            const T dn = T(deltan), c0t = T(cosa), s0t = T(sina);
            const T c = outer ? -c0t : c0t, sn = outer ? -s0t : s0t;
            const T ux = outer ? T(owx) : T(oh.wx), uy = outer ? T(owy) : T(oh.wy);
            const T ta = shelter_min(T(1), T(oh.hfa), T(ohfa), dn, c, sn, ux, uy);
            const T tw = shelter_min(T(1), T(oh.hfw), T(ohfw), dn, c, sn, T(ouw), T(ovw));
            if (ta < T(1)) atomicMin((unsigned long long*)&mina[ow], (unsigned long long)ord_key((double)ta));
            if (tw < T(1)) atomicMin((unsigned long long*)&minw[ow], (unsigned long long)ord_key((double)tw));
        }
    }
    __syncthreads();
    if (wv != 0 || !act) return;
    sha = T(ord_val(mina[lane]));
    shw = T(ord_val(minw[lane]));
    if (W2) {
        ni = xni[lane];
        qc = qc || (xfl[lane] & 1u);
        over |= xover[lane];
    }
    if (over) set_err(S, over);
    const bool qw = me.y - me.r < 0 || me.y + me.r > P.ny || me.x - me.r < 0 || me.x + me.r > P.nx;
    cand_finish<MIXED>(S, k, fs, P, no, ni, double(sha), double(shw), qc, qb, 0, 1);
    if (qc) enqueue_one(S, 0, k);
    if (qb) enqueue_one(S, 1, k);
    if (qw) enqueue_one(S, 2, k);
}

// The thread search for the dense floes whose candidates overflow the dense
// search's LDS lists (marked DENSE_OVER in ocnt, which every dense floe's
// search writes each step).  One thread per entry of the dense queue.
template <bool MIXED>
__global__ __launch_bounds__(FLOE_BLOCK) void k_cand_dense_over(DS S) {
    if (*S.err) return;
    const int k = queue_entry(S, 3);
    if (k < 0 || S.ocnt[k] != DENSE_OVER) return;
    const GeoRec me = S.geo[k];
    const CandPre pre = cand_pre(S, k);
    bool qc = false, qb = false, qw = false;
    cand_thread<MIXED, false>(S, k, me, pre, me.member, true, qc, qb, qw, nullptr);
    if (qc) enqueue_one(S, 0, k);
    if (qb) enqueue_one(S, 1, k);
    if (qw) enqueue_one(S, 2, k);
}

// ---------------------------------------------------------------------
// 3. speculative events + per-floe gather + owner state writes
// ---------------------------------------------------------------------
// Contact state of pair (i, j), owner i by storage slot, partner j by
// identity: history, last contact forces and whether it is bonded.
__device__ __forceinline__ void load_contact_state(const DS& S, int i, int j, PairState& ps) {
    ps.deltat = 0.0; ps.thetarelc = 0.0; ps.fcn = 0.0; ps.fct = 0.0;
    ps.bond = 0; ps.has_bond_rec = false;
    ps.damage = ps.deltanb = ps.deltatb = ps.trb_ji = ps.trb_ij = 0.0;
    ps.rb = ps.hb = ps.lb = ps.sb = ps.ib = 0.0;
    // both lists are searched at once, the first two records of each
    // requested together
    const int hc = S.hcnt_old[i], bc = S.bcnt_old[i];
    const int hq = find_partner<2>(S.hist_old, S.S, i, hc, j);
    const int bq = find_partner<2>(S.bond_old, S.S, i, bc, j);
    if (hq >= 0) {
        const HistRec& hr = S.hist_old[(int64_t)hq * S.S + i];
        ps.deltat = hr.deltat; ps.thetarelc = hr.thetarelc; ps.fcn = hr.fcn; ps.fct = hr.fct;
    }
    if (bq >= 0) {
        ps.has_bond_rec = true;
        ps.bond = S.bond_old[(int64_t)bq * S.S + i].bond;
    }
}

// Contact terms of the queued floes (contact_forces, plastic_contact,
// update_shape, coulomb; stepper.f90:103-146) + the owner's new histories.
// The (floe, candidate) pairs of a wavefront's floes are evaluated together,
// 64 a pass, one per lane (each lane finds its pair's floe among the lanes
// by a binary search over their offsets): a floe's inner pairs (ascending
// outer index) first, then its own candidates in distance order.  Each pass
// leaves every pair's contact terms and history outcome in LDS, and every
// floe then adds its own, in order: the same additions in the same order as
// a floe's own walk, without one lane holding the wavefront through a long
// list.
struct ContactOut {
    double fcn_c, fcr_s, fcn_s, fcr_c, r_fct, mcc, ac, fcn_ac;   // the floe's side of the terms
    double deltat, thetarelc, fcn, fct;                           // history to append (owner side)
    int32_t j;
    uint32_t kind;     // bit 0: contact terms, bit 1: history record to append
};
constexpr int CONTACT_LDS_PER_WAVE = 64 * (int)sizeof(ContactOut);

// (the mixed-precision instance gets 163 VGPRs from the looser bound, still
// 3 wavefronts per SIMD: 0.0955 -> 0.0915 ms beside the bond kernel; FP64
// would drop to 2 and keeps the bound of 3)
template <bool MIXED>
__global__ __launch_bounds__(FLOE_BLOCK, MIXED ? 2 : 3) void k_contact(DS S) {
    extern __shared__ unsigned char contact_lds[];
    ContactOut* co = reinterpret_cast<ContactOut*>(contact_lds + (threadIdx.x >> 6) * CONTACT_LDS_PER_WAVE);
    if (*S.err) return;
    const int lane = (int)(threadIdx.x & 63);
    const QueueView qv = queue_view(S, 0);
    // (the wavefront loops while any lane has an entry: its lanes share LDS
    // and shuffles)
    for (int it = 0;; ++it) {
        const int qt = queue_slot_it(qv.total, it);
        if (__ballot(qt >= 0) == 0) break;
        const int k = qt >= 0 ? queue_at(S, 0, qv, qt) : -1;
        const bool act = k >= 0;
        const int64_t SS = S.S;
        int m = 0, f = 0, fs = 0, ni = 0, no = 0;
        if (act) {
            const GeoRec me = S.geo[k];
            m = me.member;
            f = me.id;                  // identity (ordering, bond mask)
            fs = S.cval[k];             // storage slot of the floe's pair state
            ni = S.icnt[k];
            no = S.ocnt[k];
        }
        const int cnt = ni + no;
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        const int start = incl - cnt, total = __shfl(incl, 63);
        double afx = 0.0, afy = 0.0, amc = 0.0, atac = 0.0, apc = 0.0;
        int nh = 0;
        uint32_t over = 0;
        for (int c0 = 0; c0 < total; c0 += 64) {
            const int idx = c0 + lane;
            int own = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1)
                if (__shfl(start, own + step) <= idx) own += step;
            const int ko = __shfl(k, own), om = __shfl(m, own), of = __shfl(f, own);
            const int ofs = __shfl(fs, own), oni = __shfl(ni, own), ostart = __shfl(start, own);
            ContactOut r;
            r.kind = 0;
            if (idx < total) {
                const DevParams& P = S.params[om];
                const int tstep = S.members[om].tstep;
                const int first = S.members[om].first;
                const int local = idx - ostart;
                // overlap first from the position rows alone (the expression
                // kinematics evaluates, so the same decision); the full rows
                // only for contacts
                const GeoRec gf = S.geo[ko];
                if (local < oni) {
                    // pair (i, f), f the inner floe: only overlapping pairs contribute
                    const int64_t e = S.ik[(int64_t)local * SS + ko];
                    const int i = ent_id(e), qi = ent_pos(e);
                    const GeoRec gi = S.geo[qi];
                    const double dx = gi.x - gf.x, dy = gi.y - gf.y;
                    const double dn = gi.r + gf.r - sqrt(dx * dx + dy * dy);
                    if (dn > 0) {
                        Floe Fi = floe_at(S, qi);
                        Floe Fj = floe_at(S, ko);
                        Kin kn;
                        kinematics(Fi, Fj, kn);
                        PairState ps;
                        load_contact_state(S, S.cval[qi], of, ps);
                        bond_create(P, tstep, i - first, of - first, Fi, Fj, kn, ps);
                        EventOut o;
                        o.flags = 0;
                        o.ridged = -1;
                        contact_part<MIXED>(P, Fi, Fj, kn, ps, o);
                        r.kind = 1;
                        r.fcn_c = o.fcn_c; r.fcr_s = o.fcr_s; r.fcn_s = o.fcn_s; r.fcr_c = o.fcr_c;
                        r.r_fct = o.rj_fct; r.mcc = o.mcc; r.ac = o.ac; r.fcn_ac = o.fcn_ac;
                    }
                } else {
                    // pair (f, j), f the outer floe (the owner)
                    const int s = local - oni;
                    const int64_t e = S.oj[(int64_t)s * SS + ko];
                    const int j = ent_id(e), qj = ent_pos(e);
                    const GeoRec gj = S.geo[qj];
                    const double dx = gf.x - gj.x, dy = gf.y - gj.y;
                    const double dn = gf.r + gj.r - sqrt(dx * dx + dy * dy);
                    r.j = j;
                    if (!(dn > 0)) {
                        // reset_contact (reset.f90:1-15): the history goes, the
                        // last contact forces stay
                        const int q = find_partner<2>(S.hist_old, SS, ofs, S.hcnt_old[ofs], j);
                        if (q >= 0) {
                            const HistRec& h0 = S.hist_old[(int64_t)q * SS + ofs];
                            if (h0.fcn != 0.0 || h0.fct != 0.0) {
                                r.kind = 2;
                                r.deltat = 0.0; r.thetarelc = 0.0; r.fcn = h0.fcn; r.fct = h0.fct;
                                if (S.record_forces) {
                                    PairState pz;
                                    pz.fcn = h0.fcn; pz.fct = h0.fct;
                                    S.ssc[(int64_t)s * SS + ko] = stale_contact_magnitude<MIXED>(pz);
                                }
                            }
                        }
                    } else {
                        Floe Fi = floe_at(S, ko);
                        Floe Fj = floe_at(S, qj);
                        Kin kn;
                        kinematics(Fi, Fj, kn);
                        PairState ps;
                        load_contact_state(S, ofs, j, ps);
                        comp_unite(S.comp, ko, qj);      // the contact joins the floes' components
                        bond_create(P, tstep, of - first, j - first, Fi, Fj, kn, ps);
                        EventOut o;
                        o.flags = 0;
                        o.ridged = -1;
                        contact_part<MIXED>(P, Fi, Fj, kn, ps, o);
                        atomicOr(&S.evf[(int64_t)s * SS + ko], o.flags);   // (k_bond sets bits too)
                        if (S.record_forces) S.ssc[(int64_t)s * SS + ko] = o.sc;
                        r.kind = 1;
                        r.fcn_c = o.fcn_c; r.fcr_s = o.fcr_s; r.fcn_s = o.fcn_s; r.fcr_c = o.fcr_c;
                        r.r_fct = o.ri_fct; r.mcc = o.mcc; r.ac = o.ac; r.fcn_ac = o.fcn_ac;
                        if (ps.deltat != 0.0 || ps.thetarelc != 0.0 || ps.fcn != 0.0 || ps.fct != 0.0) {
                            r.kind |= 2;
                            r.deltat = ps.deltat; r.thetarelc = ps.thetarelc; r.fcn = ps.fcn; r.fct = ps.fct;
                        }
                        if (o.flags & EV_RIDGED) {
                            const int p = atomicAdd(&S.nseeds[om], 1);
                            if (p < S.SEEDK * S.members[om].n)
                                S.seeds[(int64_t)S.SEEDK * first + p] = ((uint64_t)(of - first) << 16) | (uint64_t)s;
                            else
                                over |= ERR_SEED_CAP;
                        }
                    }
                }
            }
            co[lane] = r;
            wave_lds_sync();
            // every floe adds this pass's terms of its own, in order
            if (act) {
                const int lo = max(start, c0), hi = min(start + cnt, c0 + 64);
                for (int i2 = lo; i2 < hi; ++i2) {
                    const ContactOut& t = co[i2 - c0];
                    if (t.kind & 1) {
                        if (i2 - start < ni) {
                            afx = afx + t.fcn_c - t.fcr_s;
                            afy = afy + t.fcn_s + t.fcr_c;
                            amc = amc - t.r_fct + t.mcc;
                        } else {
                            afx = afx - t.fcn_c + t.fcr_s;
                            afy = afy - t.fcn_s - t.fcr_c;
                            amc = amc - t.r_fct - t.mcc;
                        }
                        atac = atac + t.ac;
                        apc = apc - t.fcn_ac;
                    }
                    if (t.kind & 2) {
                        if (nh < S.HC) {
                            HistRec hr; hr.j = t.j; hr.pad_ = 0; hr.deltat = t.deltat; hr.thetarelc = t.thetarelc;
                            hr.fcn = t.fcn; hr.fct = t.fct;
                            S.hist_new[(int64_t)nh * SS + fs] = hr;
                            ++nh;
                        } else {
                            over |= ERR_HIST_CAP;
                        }
                    }
                }
            }
            wave_lds_sync();
        }
        if (act) {
            // contact histories of pairs that are not candidates this step
            // persist (the reference only resets a candidate pair, reset_contact)
            const int hco = S.hcnt_old[fs];
            for (int q = 0; q < hco; ++q) {
                HistRec hr = S.hist_old[(int64_t)q * SS + fs];
                if (hr.j < 0) continue;
                const bool cand = find_entry(S.oj, SS, k, no, hr.j) >= 0;
                if (!cand) {
                    if (nh < S.HC) S.hist_new[(int64_t)nh * SS + fs] = hr, ++nh;
                    else over |= ERR_HIST_CAP;
                }
            }
            S.hcnt_new[fs] = nh;
            S.acc[A_FCX * SS + k] = afx;
            S.acc[A_FCY * SS + k] = afy;
            S.acc[A_MC * SS + k] = amc;
            S.acc[A_TAC * SS + k] = atac;
            S.acc[A_PC * SS + k] = apc;
        }
        if (over) set_err(S, over);
    }
}

// One bond event of owner f (position k, storage slot fs, floe F0): pair
// (f, j) at outer slot s, written to record rq (new records: rq < 0, appended
// at nb if the event leaves one).  Returns whether the pair is bonded after
// it (the floes then have bond terms to gather).
template <bool MIXED>
__device__ __forceinline__ bool bond_event(const DS& S, const DevParams& P, int tstep, int first,
                                           int f, int k, int fs, const Floe& F0, int s, int rq,
                                           int& nb, PairState& ps, uint32_t& over) {
    const int64_t SS = S.S;
    const int64_t e = S.oj[s * SS + k];
    const int j = ent_id(e), qj = ent_pos(e);
    Floe Fi = F0;
    Floe Fj = floe_at(S, qj);
    Kin kn;
    kinematics(Fi, Fj, kn);
    bond_create(P, tstep, f - first, j - first, Fi, Fj, kn, ps);
    if (!ps.has_bond_rec) return false;
    EventOut o;
    o.flags = 0;
    o.ridged = -1;
    o.sbd = 0.0;
    bond_part<MIXED>(P, Fi, Fj, kn, ps, o);
    if (S.record_forces) S.ssb[s * SS + k] = o.sbd;
    if (rq < 0) {
        if (nb < S.BC) rq = nb++;
        else { over |= ERR_BOND_CAP; return false; }
    }
    S.bond_new[(int64_t)rq * SS + fs] = bond_from(ps, j);
    if (!(o.flags & EV_BONDED)) return false;
    BondTerms bt;
    bt.fbn_c = o.fbn_c; bt.fbt_s = o.fbt_s; bt.fbn_s = o.fbn_s; bt.fbt_c = o.fbt_c;
    bt.lri_fbt = o.lri_fbt; bt.lrj_fbt = o.lrj_fbt; bt.mbb_ji = o.mbb_ji; bt.mbb_ij = o.mbb_ij;
    bt.sb = o.sb; bt.fbn_sb = o.fbn_sb;
    S.bterm[(int64_t)rq * SS + k] = bt;          // (by the owner's position)
    atomicOr(&S.evf[s * SS + k], EV_BONDED | ((uint32_t)(rq + 1) << 16));   // (as k_contact)
    // tell the partner where its terms are: its inner slot of f
    const int qi = find_entry(S.ik, SS, qj, S.icnt[qj], f);
    if (qi >= 0) S.ibond[(int64_t)qi * SS + qj] = rq + 1;
    S.bgat[qj] = 1;
    return true;
}

// Old record b of the owner at record index rq of the new list: a record
// whose partner is a candidate this step takes part in an event (its new
// values replace it), the others carry over unchanged.
template <bool MIXED>
__device__ __forceinline__ bool bond_record(const DS& S, const DevParams& P, int tstep, int first,
                                            int f, int k, int fs, const Floe& F0, int no,
                                            const BondRec& b, int rq, int& nb, uint32_t& over) {
    const int s = find_entry(S.oj, S.S, k, no, b.j);
    if (s < 0) {
        S.bond_new[(int64_t)rq * S.S + fs] = b;
        return false;
    }
    PairState ps;
    ps.deltat = 0.0; ps.thetarelc = 0.0; ps.fcn = 0.0; ps.fct = 0.0;
    ps.has_bond_rec = true; ps.bond = b.bond; ps.damage = b.damage;
    ps.deltanb = b.deltanb; ps.deltatb = b.deltatb; ps.trb_ji = b.trb_ji; ps.trb_ij = b.trb_ij;
    ps.rb = b.rb; ps.hb = b.hb; ps.lb = b.lb; ps.sb = b.sb; ps.ib = b.ib;
    return bond_event<MIXED>(S, P, tstep, first, f, k, fs, F0, s, rq, nb, ps, over);
}

// (owner lane, record, new record index) triples of one wavefront, evaluated
// one per lane
constexpr int BOND_TCAP = 256;

// Bond events of the queued owners (bond_forces_euler, bond_breaking;
// stepper.f90:148-198): each bond is evaluated once, by its owner (the lower
// index), which writes the new record and the event's terms for both floes;
// k_final then sums every floe's terms in the reference's order.  An
// owner's records keep their order, cleared ones dropped: record q goes to
// the new index rq = the number of live records before it.  The wavefront's
// records are listed in LDS and evaluated one per lane, so that an owner with
// several bonds does not hold its lane through all of them; owners whose
// bonds form this step (step 1) append new records in candidate order and
// walk their own lists.  (3 wavefronts per SIMD, without spills: at 4 the
// kernel spilled 40 bytes, and the step, where it shares the GPU with the
// contact and wall kernels, ran 1.3 % slower.)
template <bool MIXED>
__global__ __launch_bounds__(FLOE_BLOCK, 3) void k_bond(DS S) {
    __shared__ uint32_t recs[FLOE_BLOCK / 64][BOND_TCAP];
    if (*S.err) return;
    const QueueView qv = queue_view(S, 1);
    // (the wavefront loops while any lane has an entry: its lanes share LDS
    // and shuffles)
    for (int it = 0;; ++it) {
        const int qt = queue_slot_it(qv.total, it);
        if (__ballot(qt >= 0) == 0) break;
        const int k = qt >= 0 ? queue_at(S, 1, qv, qt) : -1;
        const bool act = k >= 0;
        const int lane = (int)(threadIdx.x & 63);
        uint32_t* rl = recs[threadIdx.x >> 6];
        const int64_t SS = S.S;
        int m = 0, fs = 0, no = 0, bco = 0, fid = 0, tstep = 0, first = 0;
        bool bond_step = false;
        // partners of the first LIST_PEEK records, requested together
        int id[LIST_PEEK];
        if (act) {
            const GeoRec g = S.geo[k];
            fs = S.cval[k];
            no = S.ocnt[k];
            bco = S.bcnt_old[fs];
            // (a queued owner has at least one record)
#pragma unroll
            for (int q = 0; q < LIST_PEEK; ++q) id[q] = S.bond_old[(int64_t)min(q, max(bco - 1, 0)) * SS + fs].j;
            m = g.member;
            fid = g.id;
            // (no short-circuit tests: the loads go out together)
            const DevParams& P = S.params[m];
            const Member Mb = S.members[m];
            const int rst = P.restart, coh = P.cohesion;
            tstep = Mb.tstep;
            first = Mb.first;
            bond_step = (tstep == 1) & (rst != 1) & (coh != 0);
#pragma unroll
            for (int q = 0; q < LIST_PEEK; ++q) id[q] = q < bco ? id[q] : -1;
        }
        // the live records of the owners that form no bonds this step
        int nlive = 0;
        uint32_t over = 0;
        const bool listed = act && !bond_step;
        if (listed) {
#pragma unroll
            for (int q = 0; q < LIST_PEEK; ++q) nlive += id[q] >= 0;
            for (int q = LIST_PEEK; q < bco; ++q) nlive += S.bond_old[(int64_t)q * SS + fs].j >= 0;
        }
        int incl = nlive;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        const int total = __shfl(incl, 63);
        const bool fits = total <= BOND_TCAP;
        if (fits) {
            if (listed) {
                int at = incl - nlive, rq = 0;
                auto list_rec = [&](int q) {
                    rl[at++] = (uint32_t)lane | ((uint32_t)q << 6) | ((uint32_t)rq << 19);
                    ++rq;
                };
#pragma unroll
                for (int q = 0; q < LIST_PEEK; ++q)
                    if (id[q] >= 0) list_rec(q);    // (-1 past the count)
                for (int q = LIST_PEEK; q < bco; ++q)
                    if (S.bond_old[(int64_t)q * SS + fs].j >= 0) list_rec(q);
                S.bcnt_new[fs] = nlive;
            }
            wave_lds_sync();
            for (int c0 = 0; c0 < total; c0 += 64) {
                const int idx = c0 + lane;
                const uint32_t w = idx < total ? rl[idx] : (uint32_t)lane;
                const int own = (int)(w & 63);
                // the owner's member, step and identity, from its lane
                const int ko = __shfl(k, own), mo = __shfl(m, own), fso = __shfl(fs, own);
                const int noo = __shfl(no, own), fo = __shfl(fid, own);
                const int tso = __shfl(tstep, own), fro = __shfl(first, own);
                if (idx < total) {
                    const int q = (int)((w >> 6) & 0x1FFF), rq = (int)(w >> 19);
                    const DevParams& P = S.params[mo];
                    const Floe F0 = floe_at(S, ko);
                    const BondRec b = S.bond_old[(int64_t)q * SS + fso];
                    int nb = 0;
                    if (bond_record<MIXED>(S, P, tso, fro, fo, ko, fso, F0, noo, b, rq, nb, over))
                        S.bgat[ko] = 1;
                }
            }
        }
        if (act && (bond_step || !fits)) {
            const GeoRec me = S.geo[k];
            const DevParams& P = S.params[m];
            const int tstep = S.members[m].tstep;
            const int first = S.members[m].first;
            const int f = me.id;
            const Floe F0 = floe_at(S, k);
            bool any = false;
            int nb = 0;                     // records written so far
            for (int q = 0; q < bco; ++q) {
                const BondRec b = S.bond_old[q * SS + fs];
                if (b.j < 0) continue;
                const int rq = nb++;
                any |= bond_record<MIXED>(S, P, tstep, first, f, k, fs, F0, no, b, rq, nb, over);
            }
            // bonds that form this step (bond_create), in candidate order, appended
            if (bond_step) {
                const int nb_old = nb;
                for (int s = 0; s < no; ++s) {
                    const int j = ent_id(S.oj[s * SS + k]);
                    bool has = false;
                    for (int q = 0; q < nb_old; ++q)
                        if (S.bond_new[(int64_t)q * SS + fs].j == j) { has = true; break; }
                    if (has) continue;
                    PairState ps;
                    ps.deltat = 0.0; ps.thetarelc = 0.0; ps.fcn = 0.0; ps.fct = 0.0;
                    ps.bond = 0; ps.has_bond_rec = false;
                    ps.damage = ps.deltanb = ps.deltatb = ps.trb_ji = ps.trb_ij = 0.0;
                    ps.rb = ps.hb = ps.lb = ps.sb = ps.ib = 0.0;
                    any |= bond_event<MIXED>(S, P, tstep, first, f, k, fs, F0, s, -1, nb, ps, over);
                }
            }
            if (any) S.bgat[k] = 1;
            S.bcnt_new[fs] = nb;
        }
        if (over) set_err(S, over);
    }
}

// Bond rows of a floe with a bonded event (position k), added to a[] (rows
// at zero): the terms its owners wrote, first the pairs
// where it is the inner index (ascending outer index), then its own, in
// distance order (stepper.f90:165-198).  k_final calls it for the floes the
// repair left alone: their flags and terms are k_bond's.
__device__ __forceinline__ void gather_bonds(const DS& S, int k, int ni, int no, double* a) {
    const int64_t SS = S.S;
    // the first LIST_PEEK slots' bond words and inner partners of both sides
    // are requested together (slots past a count stand in for its last one)
    int rqi[LIST_PEEK], rqo[LIST_PEEK], pki[LIST_PEEK];
#pragma unroll
    for (int q = 0; q < LIST_PEEK; ++q) rqi[q] = rqo[q] = pki[q] = 0;
    if (ni > 0) {
#pragma unroll
        for (int q = 0; q < LIST_PEEK; ++q) {
            const int64_t r = (int64_t)min(q, ni - 1) * SS + k;
            rqi[q] = S.ibond[r];
            pki[q] = ent_pos(S.ik[r]);
        }
    }
    if (no > 0) {
#pragma unroll
        for (int q = 0; q < LIST_PEEK; ++q) rqo[q] = (int)S.evf[(int64_t)min(q, no - 1) * SS + k];
    }
#pragma unroll
    for (int q = 0; q < LIST_PEEK; ++q) {
        rqi[q] = q < ni ? rqi[q] : 0;
        rqo[q] = q < no ? rqo[q] : 0;
    }
    auto inner = [&](int pk, int rq) {
        if (rq) add_bond_j(a, S.bterm[(int64_t)(rq - 1) * SS + pk]);
    };
    auto outer = [&](uint32_t fl) {
        const int rq = (int)(fl >> 16);
        if ((fl & EV_BONDED) && rq >= 1 && rq <= S.BC) add_bond_i(a, S.bterm[(int64_t)(rq - 1) * SS + k]);
    };
#pragma unroll
    for (int q = 0; q < LIST_PEEK; ++q) inner(pki[q], rqi[q]);
    for (int q = LIST_PEEK; q < ni; ++q) inner(ent_pos(S.ik[q * SS + k]), S.ibond[q * SS + k]);
#pragma unroll
    for (int s = 0; s < LIST_PEEK; ++s) outer((uint32_t)rqo[s]);
    for (int s = LIST_PEEK; s < no; ++s) outer(S.evf[s * SS + k]);
}

// Walls of the queued floes (verify_bc, boundaries.f90:1-97).
template <bool MIXED>
__global__ __launch_bounds__(FLOE_BLOCK) void k_walls(DS S) {
    if (*S.err) return;
    // no later kernel of the step reads the build flag; k_final raises it
    // again when a floe has moved too far for the supersets
    if (blockIdx.x == 0 && threadIdx.x == 0) { S.rebuild[0] = 0; S.rebuild[1] = 0; }
    const QueueView qv = queue_view(S, 2);
    for (int it = 0;; ++it) {
        const int t = queue_slot_it(qv.total, it);
        if (t < 0) break;
        const int k = queue_at(S, 2, qv, t);
        const GeoRec me = S.geo[k];
        const int m = me.member;
        const DevParams& P = S.params[m];
        const int first = S.members[m].first;
        const int f = me.id;
        const int fs = S.cval[k];
        const int64_t SS = S.S;
        Floe Fw = floe_at(S, k);
        double hist[4] = {S.wbc[0 * SS + fs], S.wbc[1 * SS + fs], S.wbc[2 * SS + fs], S.wbc[3 * SS + fs]};
        WallOut W;
        walls<MIXED>(P, Fw, hist, W);
        S.wflag[k] = (uint8_t)((W.flags & 1u) ? WF_WALL : 0u);
        if (W.flags & 1u) S.wout[k] = W;
        if (W.flags & 4u) {
            int p = atomicAdd(&S.nseeds[m], 1);
            if (p < S.SEEDK * S.members[m].n)
                S.seeds[(int64_t)S.SEEDK * first + p] = ((uint64_t)(f - first) << 16) | WALL_SLOT;
            else
                set_err(S, ERR_SEED_CAP);
        }
    }
}

// ---------------------------------------------------------------------
// 4. ordered repair, one member per block (one wavefront)
// ---------------------------------------------------------------------
struct Heap {
    uint64_t* a;
    int n, cap;
    __device__ bool push(uint64_t k) {
        if (n >= cap) return false;
        int i = n++;
        while (i > 0) {
            int p = (i - 1) >> 1;
            if (a[p] <= k) break;
            a[i] = a[p];
            i = p;
        }
        a[i] = k;
        return true;
    }
    __device__ uint64_t pop() {
        uint64_t top = a[0];
        uint64_t last = a[--n];
        int i = 0;
        while (true) {
            int l = 2 * i + 1;
            if (l >= n) break;
            int c = (l + 1 < n && a[l + 1] < a[l]) ? l + 1 : l;
            if (a[c] >= last) break;
            a[i] = a[c];
            i = c;
        }
        if (n > 0) a[i] = last;
        return top;
    }
};

__device__ __forceinline__ void set_cur(const DS& S, int g, double r, double h, double mass,
                                        double inertia, double hfa, double hfw) {
    S.cur[0 * S.S + g] = r; S.cur[1 * S.S + g] = h; S.cur[2 * S.S + g] = mass;
    S.cur[3 * S.S + g] = inertia; S.cur[4 * S.S + g] = hfa; S.cur[5 * S.S + g] = hfw;
}

// Sets bits of storage slot f's dirty byte (the word holds three other
// floes' bytes, which other lanes may be setting; the array is allocated in
// whole words).
__device__ __forceinline__ void dirty_set(const DS& S, int f, uint32_t bits) {
    atomicOr(reinterpret_cast<uint32_t*>(S.dirty + (f & ~3)), bits << (8 * (f & 3)));
}

// Query radius of the floe at slot fs (position kf) for the events of its
// own iteration: its radius at the top of that iteration.  The first
// committed event of the iteration keeps it (cur row 6); until then no event
// of the iteration has reshaped the floe, so it is the current radius.
__device__ __forceinline__ double block_rq(const DS& S, int fs, int kf) {
    const uint8_t d = S.dirty[fs];
    return (d & DIRTY_RQ) ? S.cur[6 * S.S + fs] : ((d & DIRTY_SHAPE) ? S.cur[fs] : S.geo[kf].r);
}
__device__ __forceinline__ void keep_block_rq(const DS& S, int fs, double rq) {
    if (!(S.dirty[fs] & DIRTY_RQ)) {
        S.cur[6 * S.S + fs] = rq;
        dirty_set(S, fs, DIRTY_RQ);
    }
}

// slot of partner id g in the outer list of the floe at position kp
__device__ __forceinline__ int slot_of(const DS& S, int kp, int g) {
    return find_entry(S.oj, S.S, kp, S.ocnt[kp], g);
}

// write (or clear) owner i's new contact history entry for partner j; the
// entry lives while any of its four values is non-zero
__device__ __forceinline__ bool put_hist(const DS& S, int i, int j, const PairState& ps) {
    const bool want = ps.deltat != 0.0 || ps.thetarelc != 0.0 || ps.fcn != 0.0 || ps.fct != 0.0;
    const int n = S.hcnt_new[i];
    int q = find_partner(S.hist_new, S.S, i, n, j);
    if (want) {
        if (q < 0) {
            if (n >= S.HC) return false;
            q = n;
            S.hcnt_new[i] = n + 1;
        }
        HistRec hr; hr.j = j; hr.pad_ = 0; hr.deltat = ps.deltat; hr.thetarelc = ps.thetarelc;
        hr.fcn = ps.fcn; hr.fct = ps.fct;
        S.hist_new[(int64_t)q * S.S + i] = hr;
    } else if (q >= 0) {
        S.hist_new[(int64_t)q * S.S + i].j = -1;
    }
    return true;
}

__device__ __forceinline__ bool put_bond(const DS& S, int i, int j, bool want, const BondRec& b) {
    const int n = S.bcnt_new[i];
    int q = find_partner(S.bond_new, S.S, i, n, j);
    if (want) {
        if (q < 0) {
            if (n >= S.BC) return false;
            q = n;
            S.bcnt_new[i] = n + 1;
        }
        S.bond_new[(int64_t)q * S.S + i] = b;
    } else if (q >= 0) {
        S.bond_new[(int64_t)q * S.S + i].j = -1;
    }
    return true;
}

// Appends record r of every lane in `ma` to its owner i's list (count cnt[i],
// capacity cap) in lane order.  The lanes of one owner are adjacent in the
// batch (the keys are sorted by owner), so each takes the owner's count plus
// its rank among them, and the last of them writes the new count: no lane
// waits for another.  Every lane of the wavefront calls it.
template <typename Rec>
__device__ __forceinline__ void append_lanes(Rec* list, int32_t* cnt, int64_t SS, int cap,
                                             uint64_t ma, int i, int n, const Rec& r,
                                             uint32_t& over, uint32_t err) {
    const int lane = (int)(threadIdx.x & 63);
    const bool mine = (ma >> lane) & 1ull;
    const uint64_t below = ma & ((1ull << lane) - 1ull);
    const uint64_t above = ma & ~((2ull << lane) - 1ull);
    const int prev = below ? 63 - __clzll((long long)below) : lane;
    const int next = above ? __ffsll((long long)above) - 1 : lane;
    const int iprev = __shfl(i, prev), inext = __shfl(i, next);
    const bool head = mine && (!below || iprev != i);
    const bool last = mine && (!above || inext != i);
    const uint64_t hb = __ballot(head) & ((2ull << lane) - 1ull);
    const int start = hb ? 63 - __clzll((long long)hb) : 0;
    const int rank = __popcll(below & ~((1ull << start) - 1ull));
    if (mine) {
        const int at = n + rank;
        if (at >= cap) over |= err;
        else list[(int64_t)at * SS + i] = r;
        if (last) cnt[i] = min(at + 1, cap);
    }
}

// Lane versions of put_hist / put_bond for the batched repair: the lanes in
// `on` hold committed events, in key order.  Records found are replaced (or
// cleared) in place; new ones are appended in lane order.
__device__ __forceinline__ void put_hist_lanes(const DS& S, bool on, int i, int j,
                                               const PairState& ps, uint32_t& over) {
    const bool want = on && (ps.deltat != 0.0 || ps.thetarelc != 0.0 || ps.fcn != 0.0 || ps.fct != 0.0);
    HistRec hr;
    hr.j = j; hr.pad_ = 0; hr.deltat = ps.deltat; hr.thetarelc = ps.thetarelc; hr.fcn = ps.fcn; hr.fct = ps.fct;
    int q = -1, n = 0;
    if (on) {
        n = S.hcnt_new[i];
        q = find_partner(S.hist_new, S.S, i, n, j);
        if (q >= 0) {
            if (want) S.hist_new[(int64_t)q * S.S + i] = hr;
            else S.hist_new[(int64_t)q * S.S + i].j = -1;
        }
    }
    append_lanes(S.hist_new, S.hcnt_new, S.S, S.HC, __ballot(want && q < 0), i, n, hr, over,
                 ERR_HIST_CAP);
}

__device__ __forceinline__ void put_bond_lanes(const DS& S, bool on, int i, int j, bool want,
                                               const BondRec& b, uint32_t& over) {
    want = want && on;
    int q = -1, n = 0;
    if (on) {
        n = S.bcnt_new[i];
        q = find_partner(S.bond_new, S.S, i, n, j);
        if (q >= 0) {
            if (want) S.bond_new[(int64_t)q * S.S + i] = b;
            else S.bond_new[(int64_t)q * S.S + i].j = -1;
        }
    }
    append_lanes(S.bond_new, S.bcnt_new, S.S, S.BC, __ballot(want && q < 0), i, n, b, over,
                 ERR_BOND_CAP);
}

// Progress of one member's repair (the same in every lane).
struct RepairState {
    int64_t last;        // key of the last event recomputed (-1: none)
    int nt, nwt;         // repair table entries used
    int64_t repaired;
};

// Every later event of floe g, reshaped by event (f, s), goes onto the heap:
// the pairs (i2, g) with f < i2 < g, then g's own events after s (all of
// them if g != f) and g's walls.
__device__ __forceinline__ void push_later_events(const DS& S, Heap& H, int first, int f, int s,
                                                  int g, int kg, uint32_t& over) {
    const int64_t SS = S.S;
    const int gb = g - first;
    if (g != f) {
        const int gi = S.icnt[kg];
        for (int q = 0; q < gi; ++q) {
            const int64_t ei = S.ik[(int64_t)q * SS + kg];
            const int i2 = ent_id(ei);
            if (i2 <= f) continue;
            const int sk = slot_of(S, ent_pos(ei), g);
            if (sk >= 0 && !H.push(((uint64_t)(i2 - first) << 16) | (uint64_t)sk)) over |= ERR_REPAIR_CAP;
        }
    }
    const int s0 = (g == f) ? s + 1 : 0;
    const int gno = S.ocnt[kg];
    for (int s2 = s0; s2 < gno; ++s2)
        if (!H.push(((uint64_t)gb << 16) | (uint64_t)s2)) over |= ERR_REPAIR_CAP;
    if (!H.push(((uint64_t)gb << 16) | WALL_SLOT)) over |= ERR_REPAIR_CAP;
}

// Every contact floe (queue 0) points at the root of its contact component
// (no unions run any more).  Every floe of a contact is in the queue.
__global__ __launch_bounds__(FLOE_BLOCK) void k_comp(DS S) {
    if (*S.err) return;
    const QueueView qv = queue_view(S, 0);
    for (int it = 0;; ++it) {
        const int t = queue_slot_it(qv.total, it);
        if (t < 0) break;
        const int k = queue_at(S, 0, qv, t);
        S.comp[k] = comp_find(S.comp, k);
    }
}

// One event at a time off the binary heap in HBM (single lane).
template <bool MIXED>
__device__ void repair_serial(const DS& S, const DevParams& P, int first, int nmem, int tstep,
                              Heap& H, EventOut* etab, int tcap, WallOut* wtab, RepairState& st,
                              uint32_t& over) {
    const int64_t SS = S.S;
    while (H.n > 0 && !over) {
        const uint64_t key = H.pop();
        if ((int64_t)key == st.last) continue;
        st.last = (int64_t)key;
        const int b = (int)(key >> 16), s = (int)(key & 0xFFFF);
        // keys name an event of this member (a broken key is an internal
        // error, never an out-of-range access)
        if (key >> 16 >= (uint64_t)nmem || (s != WALL_SLOT && s >= S.CO)) { over |= ERR_REPAIR_CAP; break; }
        const int f = first + b;              // identity
        const int fs = S.sid_of[f];           // storage slot
        const int kf = S.pos_id[f];           // position (loaded beside the slot)
        ++st.repaired;
        if (s == WALL_SLOT) {
            Floe F = load_floe_cur(S, fs, kf);
            double hist[4] = {S.wbc[0 * SS + fs], S.wbc[1 * SS + fs], S.wbc[2 * SS + fs], S.wbc[3 * SS + fs]};
            WallOut W;
            walls<MIXED>(P, F, hist, W);
            if (st.nwt >= nmem) { over |= ERR_REPAIR_CAP; break; }
            wtab[st.nwt] = W;
            S.ovw[kf] = st.nwt++;
            wflag_set(S, kf, WF_REPAIR_WALL | WF_FIXUP);
            if (W.flags & 4u) {
                set_cur(S, fs, W.nr, W.nh, W.nmass, W.ninertia, W.nhfa, W.nhfw);
                dirty_set(S, fs, DIRTY_SHAPE);
            }
            continue;
        }
        const double rq = block_rq(S, fs, kf);   // query radius at the top of iteration f
        keep_block_rq(S, fs, rq);
        const int64_t e = S.oj[(int64_t)s * SS + kf];
        const int j = ent_id(e);
        const int js = S.cval[ent_pos(e)];
        Floe Fi = load_floe_cur(S, fs, kf);
        Floe Fj = load_floe_cur(S, js, ent_pos(e));
        wflag_set(S, kf, WF_FIXUP);
        wflag_set(S, ent_pos(e), WF_FIXUP);
        EventOut o;
        const double dx = Fi.x - Fj.x, dy = Fi.y - Fj.y;
        const double d = sqrt(dx * dx + dy * dy);
        PairState ps;
        load_pair_state(S, fs, j, ps);
        if (!(d <= rq + P.rtree)) {
            // no longer a candidate: the pair keeps its old state
            o.flags = EV_REMOVED;
            o.ridged = -1;
        } else {
            const double2 wf = wind_at(S, kf);
            pair_event<MIXED>(P, tstep, b, j - first, Fi, Fj, wf.x, wf.y, P.uw, P.vw, ps, o, 0);
        }
        if (!put_hist(S, fs, j, ps)) over |= ERR_HIST_CAP;
        if (!put_bond(S, fs, j, ps.has_bond_rec, bond_from(ps, j))) over |= ERR_BOND_CAP;
        if (o.flags & EV_RIDGED) {
            const int g = o.ridged == 0 ? f : j;
            const int gs = o.ridged == 0 ? fs : js;
            const int kg = o.ridged == 0 ? kf : ent_pos(e);
            set_cur(S, gs, o.nr, o.nh, o.nmass, o.ninertia, o.nhfa, o.nhfw);
            if (!(S.dirty[gs] & DIRTY_SHAPE)) {
                dirty_set(S, gs, DIRTY_SHAPE);
                push_later_events(S, H, first, f, s, g, kg, over);
            }
        }
        if (st.nt >= tcap) { over |= ERR_REPAIR_CAP; break; }
        etab[st.nt] = o;
        S.ov[(int64_t)s * SS + kf] = st.nt++;
        S.evf[(int64_t)s * SS + kf] = o.flags | EV_TABLE;
    }
}

// number of a[0..n) below v (strictly, or at most v)
__device__ __forceinline__ int count_below(const int64_t* a, int n, int64_t v, bool or_equal) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const bool left = or_equal ? a[mid] <= v : a[mid] < v;
        if (left) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// The recomputation follows the reference's loop order, key
// (f - first) << 16 | slot.  Between two reshapes the recomputed events are
// independent: an event reads the radius, thickness, mass, inertia and
// drafts of its floes, which only a reshape changes, and writes the state of
// its own pair.  So the wavefront evaluates the next `repair_lanes` pending
// keys at once, one per lane, and commits them in key order up to and
// including the first that reshapes a floe; the rest stay pending and are
// evaluated again with the new state.  The pending keys are a sorted array in
// LDS; those a reshape adds (every later event of the reshaped floe, produced
// in order) are merged in.  A member whose pending keys outgrow `repair_cap`
// continues one event at a time on its binary heap in HBM.
constexpr int REPAIR_PCAP = 1024;    // pending keys held in LDS
constexpr int REPAIR_NCAP = 512;     // keys one reshape adds, at most

template <bool MIXED>
__global__ __launch_bounds__(64) void k_repair(DS S, int nmembers) {
    __shared__ int64_t pk[2][REPAIR_PCAP];
    __shared__ int64_t nk[REPAIR_NCAP];
    // each lane's result until the round decides whether it commits (one
    // EventOut-sized slot per lane; a wall lane's WallOut uses its start)
    __shared__ uint64_t stage_buf[64 * sizeof(EventOut) / 8];
    __shared__ int tcomp[64];      // components of a round's reshaping lanes so far
    // the member's parameters, copied to LDS: the physics reads them at LDS
    // latency (this wavefront is the only one working on the member)
    __shared__ uint64_t pbuf[sizeof(DevParams) / 8];
    static_assert(sizeof(DevParams) % 8 == 0 && sizeof(DevParams) / 8 <= 64, "DevParams copied by lane");
    const int m = blockIdx.x;
    if (m >= nmembers || *S.err) return;
    if (!member_live(S.members[m])) return;
    int ns = S.nseeds[m];
    if (ns == 0) return;
    if ((int)threadIdx.x < (int)(sizeof(DevParams) / 8))
        pbuf[threadIdx.x] = reinterpret_cast<const uint64_t*>(S.params + m)[threadIdx.x];
    __syncthreads();
    const DevParams& P = *reinterpret_cast<const DevParams*>(pbuf);
    const int first = S.members[m].first;
    const int nmem = S.members[m].n;
    const int tstep = S.members[m].tstep;
    const int64_t SS = S.S;
    if (ns > S.SEEDK * nmem) ns = S.SEEDK * nmem;
    const int lane = (int)threadIdx.x;
    const uint64_t lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    const uint64_t* seeds = S.seeds + (int64_t)S.SEEDK * first;
    EventOut* etab = S.etab + (int64_t)S.TABK * first;
    const int tcap = S.TABK * nmem;
    WallOut* wtab = S.wtab + first;
    RepairState st;
    st.last = -1; st.nt = 0; st.nwt = 0; st.repaired = 0;
    uint32_t over = 0;
    Heap H;
    H.a = S.heap + (int64_t)S.HEAPK * first;
    H.n = 0;
    H.cap = S.HEAPK * nmem;
    EventOut* stage_e = reinterpret_cast<EventOut*>(stage_buf);
    WallOut& stage_w_mine = *reinterpret_cast<WallOut*>(stage_e + threadIdx.x);
    const int lanes = min(max(S.repair_lanes, 1), 64);
    const int cap = min(max(S.repair_cap, 1), REPAIR_PCAP);
    bool serial = lanes == 1 || ns > cap;
    int cb = 0, head = 0, np = 0;
    if (serial) {
        if (lane == 0)
            for (int p = 0; p < ns; ++p)
                if (!H.push(seeds[p])) over |= ERR_REPAIR_CAP;
    } else {
        const int p2 = pow2_at_least(ns);
        for (int i = lane; i < p2; i += 64) pk[0][i] = i < ns ? (int64_t)seeds[i] : INT64_MAX;
        __syncthreads();
        bitonic_inner(pk[0], p2, lane);
        np = ns;
    }
    while (!serial && np > 0) {
        const int nb = min(np, lanes);
        const int64_t key = lane < nb ? pk[cb][head + lane] : INT64_MAX;
        const int64_t prev = lane == 0 ? st.last : (lane < nb ? pk[cb][head + lane - 1] : INT64_MAX);
        bool act = lane < nb && key != prev;            // repeated keys are done once
        // keys name an event of this member (a broken key is an internal
        // error, never an out-of-range access)
        if (act && (key < 0 || (key >> 16) >= nmem || ((key & 0xFFFF) != WALL_SLOT && (key & 0xFFFF) >= S.CO))) {
            over |= ERR_REPAIR_CAP;
            act = false;
        }
        if (__ballot(over != 0)) break;
        const int64_t kk = act ? key : 0;
        const int b = (int)(kk >> 16), s = (int)(kk & 0xFFFF);
        const int f = first + b;
        const bool wall = act && s == WALL_SLOT;
        const bool pair = act && s != WALL_SLOT;
        int fs = 0, kf = 0, j = 0, js = 0, kj = 0, ridged = -1, ca = -1, cbm = -1;
        bool reshape = false;
        uint32_t oflags = 0;
        double rql = 0.0, nr = 0.0, nh = 0.0, nmass = 0.0, ninertia = 0.0, nhfa = 0.0, nhfw = 0.0;
        PairState ps;
        if (act) {
            fs = S.sid_of[f];
            kf = S.pos_id[f];
        }
        if (wall) {
            ca = cbm = S.comp[kf];
            Floe F = load_floe_cur(S, fs, kf);
            double hist[4] = {S.wbc[0 * SS + fs], S.wbc[1 * SS + fs], S.wbc[2 * SS + fs], S.wbc[3 * SS + fs]};
            WallOut W;
            walls<MIXED>(P, F, hist, W);
            stage_w_mine = W;
            reshape = (W.flags & 4u) != 0;
            nr = W.nr; nh = W.nh; nmass = W.nmass; ninertia = W.ninertia; nhfa = W.nhfa; nhfw = W.nhfw;
        }
        if (pair) {
            rql = block_rq(S, fs, kf);
            const int64_t e = S.oj[(int64_t)s * SS + kf];
            j = ent_id(e);
            kj = ent_pos(e);
            js = S.cval[kj];
            ca = S.comp[kf];
            cbm = S.comp[kj];
            Floe Fi = load_floe_cur(S, fs, kf);
            Floe Fj = load_floe_cur(S, js, kj);
            const double dx = Fi.x - Fj.x, dy = Fi.y - Fj.y;
            const double d = sqrt(dx * dx + dy * dy);
            load_pair_state(S, fs, j, ps);
            EventOut o;
            if (!(d <= rql + P.rtree)) {
                o.flags = EV_REMOVED;
                o.ridged = -1;
            } else {
                const double2 wf = wind_at(S, kf);
                pair_event<MIXED>(P, tstep, b, j - first, Fi, Fj, wf.x, wf.y, P.uw, P.vw, ps, o, 0);
            }
            stage_e[lane] = o;
            oflags = o.flags;
            reshape = (o.flags & EV_RIDGED) != 0;
            ridged = o.ridged;
            nr = o.nr; nh = o.nh; nmass = o.nmass; ninertia = o.ninertia; nhfa = o.nhfa; nhfw = o.nhfw;
        }
        // Commit every event of the batch whose floes no uncommitted event
        // before it (in key order) can still change, and whose reshape no
        // such event can still read.  A floe changes only when an event of it
        // ridges, so only through contacts: within its contact component
        // (comp, from the step's contacts at its start radii; radii only
        // shrink).  An event before a lane that is not committed yet (a
        // deferred lane, one a reshape adds, one those add in turn) belongs
        // to a component of T, the components of the reshaping lanes before
        // it (a lane waits when a floe of it lies in T, so its own component
        // is in T if it ridges at all), or it reads floes in T and their
        // candidate neighbours.  So, going through the reshaping lanes in
        // order: the component of each joins T, and the later lanes in T
        // wait; a reshaping lane itself waits if a waiting lane before it
        // reads the floe it reshapes or a candidate neighbour of that floe
        // lies in T.  (A lane that waits ridges nothing new: if its floes
        // touch, they share a component in T; if not, it cannot ridge.)
        uint64_t defer = 0, done = 0;
        int nT = 0;
        const uint64_t rsh = __ballot(reshape);
        const int kgl = (wall || ridged == 0) ? kf : kj;     // the floe a reshaping lane reshapes
        for (;;) {
            const uint64_t t = rsh & ~defer & ~done;
            if (!t) break;
            const int cl = __ffsll((long long)t) - 1;
            done |= 1ull << cl;
            const int cc = __shfl(ca, cl);
            const int kgc = __shfl(kgl, cl);
            bool exposed = __ballot(lane < cl && ((defer >> lane) & 1ull) && (kf == kgc || (pair && kj == kgc))) != 0;
            if (!exposed && nT > 0) {
                const int no_g = S.ocnt[kgc], ncand = no_g + S.icnt[kgc];
                bool hit = false;
                for (int q0 = 0; q0 < ncand; q0 += 64) {
                    const int q = q0 + lane;
                    if (q < ncand) {
                        const int64_t e = q < no_g ? S.oj[(int64_t)q * SS + kgc] : S.ik[(int64_t)(q - no_g) * SS + kgc];
                        const int cn = S.comp[ent_pos(e)];
                        for (int u = 0; u < nT; ++u) hit = hit || cn == tcomp[u];
                    }
                    if (__ballot(hit)) break;
                }
                exposed = __ballot(hit) != 0;
            }
            if (exposed) defer |= 1ull << cl;
            if (lane == 0) tcomp[nT] = cc;
            ++nT;
            wave_lds_sync();
            defer |= __ballot(act && lane > cl && (ca == cc || cbm == cc));
        }
        const bool com = act && !((defer >> lane) & 1ull);
        const uint64_t mce = __ballot(com && pair), mcw = __ballot(com && wall);
        const int rank_e = __popcll(mce & lt), rank_w = __popcll(mcw & lt);
        if (st.nt + __popcll(mce) > tcap || st.nwt + __popcll(mcw) > nmem) { over |= ERR_REPAIR_CAP; break; }
        if (com && pair) {
            const int at = st.nt + rank_e;
            etab[at] = stage_e[lane];
            wflag_set(S, kf, WF_FIXUP);
            wflag_set(S, kj, WF_FIXUP);
            S.ov[(int64_t)s * SS + kf] = at;
            S.evf[(int64_t)s * SS + kf] = oflags | EV_TABLE;
            keep_block_rq(S, fs, rql);
        }
        if (com && wall) {
            const int at = st.nwt + rank_w;
            wtab[at] = stage_w_mine;
            S.ovw[kf] = at;
            wflag_set(S, kf, WF_REPAIR_WALL | WF_FIXUP);
        }
        put_hist_lanes(S, com && pair, fs, j, ps, over);
        put_bond_lanes(S, com && pair, fs, j, ps.has_bond_rec, bond_from(ps, j), over);
        int g = 0, gs = 0, kg = 0;
        bool adds = false;
        if (com && reshape) {
            g = (wall || ridged == 0) ? f : j;
            gs = (wall || ridged == 0) ? fs : js;
            kg = (wall || ridged == 0) ? kf : kj;
            set_cur(S, gs, nr, nh, nmass, ninertia, nhfa, nhfw);
            // a wall is a floe's last event: nothing of it follows
            adds = !wall && !(S.dirty[gs] & DIRTY_SHAPE);
            dirty_set(S, gs, DIRTY_SHAPE);
        }
        st.nt += __popcll(mce);
        st.nwt += __popcll(mcw);
        st.repaired += __popcll(mce | mcw);
        // the next round starts with the deferred keys, in order, then the
        // keys after the batch; one equal to the batch's last key is dropped
        // if that key is done (its first lane committed, or it repeats the
        // key done before the batch)
        const int64_t lastkey = __shfl(key, nb - 1);
        const uint64_t same = __ballot(lane < nb && key == lastkey);
        const int fl = __ffsll((long long)same) - 1;
        // (a first lane that is not active repeats the key done before)
        const bool last_done = __shfl((int)act, fl) ? !((defer >> fl) & 1ull) : true;
        const int nd = __popcll(defer);
        int skip = 0;
        if (last_done)
            while (nb + skip < np && pk[cb][head + nb + skip] == lastkey) ++skip;
        st.last = last_done ? lastkey : -1;
        wave_lds_sync();
        // deferred keys go just before the keys after the batch
        const int nhead = head + nb + skip - nd;
        if ((defer >> lane) & 1ull) pk[cb][nhead + __popcll(defer & lt)] = key;
        wave_lds_sync();
        if (__ballot(over != 0)) break;
        head = nhead;
        np = np - nb - skip + nd;
        if (!__ballot(adds)) continue;
        // each committed reshape (of a floe not reshaped before) adds every
        // later event of its floe g: one sorted run per reshape, the runs
        // sorted together when there are several
        const uint64_t madd = __ballot(adds);
        const int rest = np;
        int K = 0;
        bool spill_heap = false;
        for (uint64_t ma = madd; ma; ma &= ma - 1) {
            const int cl = __ffsll((long long)ma) - 1;
            const int gc = __shfl(g, cl), fc = __shfl(f, cl), kgc = __shfl(kg, cl);
            const int sc = __shfl(s, cl);
            const int gi = gc != fc ? S.icnt[kgc] : 0;
            const int s0 = gc == fc ? sc + 1 : 0;
            const int gno = S.ocnt[kgc];
            const int kmax = gi + max(gno - s0, 0) + 1;
            if (K + kmax > REPAIR_NCAP || rest + K + kmax > cap) { spill_heap = true; break; }
            for (int q0 = 0; q0 < gi; q0 += 64) {
                const int q = q0 + lane;
                bool ok = false;
                int64_t nkey = 0;
                if (q < gi) {
                    const int64_t ei = S.ik[(int64_t)q * SS + kgc];
                    const int i2 = ent_id(ei);
                    if (i2 > fc) {
                        const int sk = slot_of(S, ent_pos(ei), gc);
                        if (sk >= 0) { ok = true; nkey = ((int64_t)(i2 - first) << 16) | sk; }
                    }
                }
                const uint64_t mk = __ballot(ok);
                if (ok) nk[K + __popcll(mk & lt)] = nkey;
                K += __popcll(mk);
            }
            for (int s2 = s0 + lane; s2 < gno; s2 += 64) nk[K + s2 - s0] = ((int64_t)(gc - first) << 16) | s2;
            K += max(gno - s0, 0);
            if (lane == 0) nk[K] = ((int64_t)(gc - first) << 16) | WALL_SLOT;
            K += 1;
        }
        if (spill_heap) {
            // too many pending keys for LDS: on to the heap in HBM
            if (lane == 0) {
                for (int i = 0; i < rest; ++i)
                    if (!H.push((uint64_t)pk[cb][head + i])) over |= ERR_REPAIR_CAP;
            }
            for (uint64_t ma = madd; ma; ma &= ma - 1) {
                const int cl = __ffsll((long long)ma) - 1;
                const int gc = __shfl(g, cl), fc = __shfl(f, cl), kgc = __shfl(kg, cl);
                const int sc = __shfl(s, cl);
                if (lane == 0) push_later_events(S, H, first, fc, sc, gc, kgc, over);
            }
            serial = true;
            break;
        }
        wave_lds_sync();
        if (__popcll(madd) > 1) {
            const int p2 = pow2_at_least(K);
            for (int i = K + lane; i < p2; i += 64) nk[i] = INT64_MAX;
            wave_lds_sync();
            bitonic_inner(nk, p2, lane);
        }
        // merge the two sorted runs (pending keys first among equals)
        const int64_t* A = &pk[cb][head];
        int64_t* O = pk[cb ^ 1];
        for (int i = lane; i < rest; i += 64) O[i + count_below(nk, K, A[i], false)] = A[i];
        for (int i = lane; i < K; i += 64) O[i + count_below(A, rest, nk[i], true)] = nk[i];
        wave_lds_sync();
        cb ^= 1;
        head = 0;
        np = rest + K;
    }
    if (serial && lane == 0)
        repair_serial<MIXED>(S, P, first, nmem, tstep, H, etab, tcap, wtab, st, over);
    if (lane == 0) {
        S.netab[m] = st.nt;
        S.nwtab[m] = st.nwt;
        S.stats[4 * m + 2] = ns;
        S.stats[4 * m + 3] = st.repaired;
    }
    if (over) set_err(S, over);
}

// ---------------------------------------------------------------------
// 5. re-gather floes touched by the repair
// ---------------------------------------------------------------------
// Events of one wavefront's touched floes evaluated together (LDS per
// wavefront: one pass of 64 results).
constexpr int FIX_LDS_PER_WAVE = 64 * (int)sizeof(EventOut);
static_assert(sizeof(EventOut) % 8 == 0, "EventOut is copied in 8-byte words");

__device__ __forceinline__ void copy_event(EventOut* dst, const EventOut* src) {
    const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
    uint64_t* d = reinterpret_cast<uint64_t*>(dst);
#pragma unroll
    for (int w = 0; w < (int)(sizeof(EventOut) / 8); ++w) d[w] = s[w];
}

// The floes the repair touched re-gather every event of theirs, in the
// reference's order (inner pairs by ascending outer index, then their own
// candidates in distance order): the repair's result where the table holds
// one, else the event evaluated again (its inputs are final).  All the
// wavefront's events are evaluated together, 64 a pass (each lane finds its
// event's floe among the lanes), and after each pass every floe adds its own
// in order, so that a floe with many candidates does not hold one lane
// through all of them.
template <bool MIXED>
__global__ __launch_bounds__(FLOE_BLOCK, 2) void k_fixup(DS S) {
    extern __shared__ unsigned char fix_lds[];
    EventOut* ev = reinterpret_cast<EventOut*>(fix_lds + (threadIdx.x >> 6) * FIX_LDS_PER_WAVE);
    // The floes the repair touched (the fixup queue, filled by k_fixq), a few
    // per wavefront: every wavefront of the grid (about one resident round)
    // takes F consecutive entries, F as small as the queue allows, so that
    // the floes' re-gathers run side by side rather than in a few long
    // wavefronts.
    const int lane = (int)(threadIdx.x & 63);
    const QueueView qv = queue_view(S, 4);
    const int nw = (int)(gridDim.x * blockDim.x) >> 6;
    const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int F = min(64, max(1, (qv.total + nw - 1) / nw));
    for (int t0 = w * F; t0 < qv.total; t0 += nw * F) {     // (the same for all lanes)
        const int t = t0 + lane;
        bool act = lane < F && t < qv.total && !*S.err;
        const int k = act ? queue_at(S, 4, qv, t) : 0;
        int m = 0;
        if (act) {
            m = S.geo[k].member;
            act = m >= 0;
        }
        if (!__ballot(act)) continue;
        if (!act) m = 0;
        const int f = act ? S.geo[k].id : 0;
        const int fs = act ? (int)S.cval[k] : 0;
        const int ni = act ? S.icnt[k] : 0, no = act ? S.ocnt[k] : 0;
        const int cnt = ni + no;
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        const int start = incl - cnt, total = __shfl(incl, 63);
        const int64_t SS = S.S;
        double acc[NACC];
        acc_init(acc);
        for (int c0 = 0; c0 < total; c0 += 64) {
            const int idx = c0 + lane;
            int own = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1)
                if (__shfl(start, own + step) <= idx) own += step;
            const int ko = __shfl(k, own), om = __shfl(m, own), of = __shfl(f, own);
            const int ofs = __shfl(fs, own), oni = __shfl(ni, own), ostart = __shfl(start, own);
            if (idx < total) {
                const DevParams& P = S.params[om];
                const int tstep = S.members[om].tstep;
                const int first = S.members[om].first;
                const EventOut* etab = S.etab + (int64_t)S.TABK * first;
                const int local = idx - ostart;
                const bool inner = local < oni;
                int tab = -1, qp = 0, pid = 0;
                if (inner) {
                    const int64_t e = S.ik[(int64_t)local * SS + ko];
                    pid = ent_id(e);
                    qp = ent_pos(e);
                    const int sk = slot_of(S, qp, of);
                    tab = sk >= 0 && (S.evf[(int64_t)sk * SS + qp] & EV_TABLE) ? S.ov[(int64_t)sk * SS + qp] : -1;
                } else {
                    const int s2 = local - oni;
                    tab = (S.evf[(int64_t)s2 * SS + ko] & EV_TABLE) ? S.ov[(int64_t)s2 * SS + ko] : -1;
                    const int64_t e = S.oj[(int64_t)s2 * SS + ko];
                    pid = ent_id(e);
                    qp = ent_pos(e);
                }
                if (tab >= 0) {
                    copy_event(&ev[lane], &etab[tab]);
                    if (ev[lane].flags & EV_REMOVED) ev[lane].flags = 0;   // adds nothing
                } else {
                    // inner: pair (i, f) with i's wind; outer: pair (f, j) with f's wind
                    PairState ps;
                    load_pair_state<1>(S, inner ? (int)S.cval[qp] : ofs, inner ? of : pid, ps);
                    const double2 w = wind_at(S, inner ? qp : ko);
                    Floe Fa = floe_at(S, inner ? qp : ko), Fb = floe_at(S, inner ? ko : qp);
                    EventOut o;
                    pair_event<MIXED>(P, tstep, (inner ? pid : of) - first, (inner ? of : pid) - first,
                                      Fa, Fb, w.x, w.y, P.uw, P.vw, ps, o, inner ? 2 : 1);
                    copy_event(&ev[lane], &o);
                }
            }
            wave_lds_sync();
            // every floe adds this pass's events of its own, in order
            if (act) {
                const int lo = max(start, c0), hi = min(start + cnt, c0 + 64);
                for (int i2 = lo; i2 < hi; ++i2) {
                    if (i2 - start < ni) apply_j(acc, ev[i2 - c0]);
                    else apply_i(acc, ev[i2 - c0]);
                }
            }
            wave_lds_sync();
        }
        if (act) {
#pragma unroll
            for (int a = 0; a < NACC; ++a) S.acc[a * SS + k] = acc[a];
            S.cls[k] = CLS_CONTACT | CLS_BOND;
        }
    }
}

// The positions the repair touched (WF_FIXUP) into the fixup queue, one
// atomic per wavefront; the grid is laid out as the candidate search's.
__global__ __launch_bounds__(FLOE_BLOCK) void k_fixq(DS S) {
    const int k = xcd_block() * blockDim.x + threadIdx.x;
    const bool want = k < S.N && !*S.err && (S.wflag[k] & WF_FIXUP);
    enqueue(S, want, 4, k);
}

// ---------------------------------------------------------------------
// 5b. stress diagnostics (stepper.f90:238-271), only when forces are
//     recorded: for every candidate pair (j,i), both floes take
//     sig -= S * (direction terms) * r(i) with S = |(fcn,fct)| + |(fbn,fbt)|,
//     the pair's current or last contact forces and current bond forces, in
//     the reference's per-floe order.
// ---------------------------------------------------------------------
__device__ __forceinline__ void stress_terms(double Sv, double c, double sn, double ri, double* a) {
    a[A_SXX] = a[A_SXX] - Sv * c * ri * c;
    a[A_SYY] = a[A_SYY] - Sv * sn * ri * sn;
    a[A_SXY] = a[A_SXY] - Sv * sn * ri * c;
    a[A_SYX] = a[A_SYX] - Sv * c * ri * sn;
}

__global__ __launch_bounds__(FLOE_BLOCK) void k_stress(DS S) {
    const int k = xcd_block() * blockDim.x + threadIdx.x;
    if (k >= S.N || *S.err) return;
    const GeoRec me = S.geo[k];
    const int m = me.member;
    if (m < 0) return;
    const int f = me.id;
    const int64_t SS = S.S;
    const EventOut* etab = S.etab + (int64_t)S.TABK * S.members[m].first;
    double a[NACC];
    a[A_SXX] = a[A_SYY] = a[A_SXY] = a[A_SYX] = 0.0;
    // pairs (i, f), f inner, ascending i
    const int ni = S.icnt[k];
    for (int q = 0; q < ni; ++q) {
        const int64_t e = S.ik[q * SS + k];
        const int qi = ent_pos(e);
        const int s = slot_of(S, qi, f);
        if (s < 0) continue;
        const int idx = (S.evf[(int64_t)s * SS + qi] & EV_TABLE) ? S.ov[(int64_t)s * SS + qi] : -1;
        const GeoRec gi = S.geo[qi];
        double sc, sb, ri;
        if (idx >= 0) {
            const EventOut& o = etab[idx];
            if (o.flags & EV_REMOVED) continue;
            sc = o.sc; sb = o.sbd; ri = o.sri;
        } else {
            sc = S.ssc[(int64_t)s * SS + qi]; sb = S.ssb[(int64_t)s * SS + qi]; ri = gi.r;
        }
        const double dx = gi.x - me.x, dy = gi.y - me.y;
        const double dist = sqrt(dx * dx + dy * dy);
        const double c = (me.x - gi.x) / dist, sn = (me.y - gi.y) / dist;
        stress_terms(sc + sb, c, sn, ri, a);
    }
    // pairs (f, j), f outer, distance order
    const int no = S.ocnt[k];
    for (int s = 0; s < no; ++s) {
        const int64_t e = S.oj[s * SS + k];
        const int qj = ent_pos(e);
        const int idx = (S.evf[s * SS + k] & EV_TABLE) ? S.ov[s * SS + k] : -1;
        double sc, sb, ri;
        if (idx >= 0) {
            const EventOut& o = etab[idx];
            if (o.flags & EV_REMOVED) continue;
            sc = o.sc; sb = o.sbd; ri = o.sri;
        } else {
            sc = S.ssc[s * SS + k]; sb = S.ssb[s * SS + k]; ri = me.r;
        }
        const GeoRec gj = S.geo[qj];
        const double dx = me.x - gj.x, dy = me.y - gj.y;
        const double dist = sqrt(dx * dx + dy * dy);
        const double c = (gj.x - me.x) / dist, sn = (gj.y - me.y) / dist;
        stress_terms(sc + sb, c, sn, ri, a);
    }
    S.acc[A_SXX * SS + k] = a[A_SXX];
    S.acc[A_SYY * SS + k] = a[A_SYY];
    S.acc[A_SXY * SS + k] = a[A_SXY];
    S.acc[A_SYX * SS + k] = a[A_SYX];
}

// ---------------------------------------------------------------------
// 6. final properties, forcing + coriolis, totals, integration
//    (stepper.f90:295-343, integration.f90:1-38)
// ---------------------------------------------------------------------
// (REC: forces are recorded; the other instance carries no code for the
// recorded columns and stress rows.  The mixed-precision instance without
// them fits 5 wavefronts per SIMD in 96 VGPRs, one value kept in scratch
// from the first loads to the writes: its chains of dependent loads, the
// gridded wind's among them, gain more from the extra wavefront)
template <bool MIXED, bool REC>
__global__ __launch_bounds__(FLOE_BLOCK, (MIXED && !REC) ? 5 : 4) void k_final(DS S) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S.N) return;
    // the position's rows are requested at once, before anything is tested
    const GeoRec me = S.geo[k];
    const int f = S.cval[k];        // storage slot
    Floe F = floe_at(S, k);
    const uint8_t wf = S.wflag[k], cl = S.cls[k], bg = S.bgat[k];
    const int ni = S.icnt[k], no = S.ocnt[k];
    if (*S.err) return;
    // every earlier kernel of the step has finished: the step is complete
    if (k == 0) *S.steps_done += 1;
    const int m = me.member;
    if (m < 0) return;
    const DevParams& P = S.params[m];
    const int64_t SS = S.S;
    const uint8_t dflags = S.dirty[f];
    const bool dirty = (dflags & DIRTY_SHAPE) != 0;
    if (dirty) {
        F.r = S.cur[0 * SS + f]; F.h = S.cur[1 * SS + f]; F.mass = S.cur[2 * SS + f];
        F.inertia = S.cur[3 * SS + f]; F.hfa = S.cur[4 * SS + f]; F.hfw = S.cur[5 * SS + f];
    }
    WallOut W;
    if (wf & WF_REPAIR_WALL) {
        W = S.wtab[S.members[m].first + S.ovw[k]];   // recomputed by the repair
    } else if (wf & WF_WALL) {
        W = S.wout[k];
    } else {
        // no wall touched: reset_boundary(i, 0) cleared every wall history
        W.fx_bc = W.fy_bc = W.m_bc = W.ta_bc = W.p_bc = 0.0;
        W.deltat_bc1 = W.deltat_bc2 = W.theta_bc1 = W.theta_bc2 = 0.0;
        W.sxx_bc = W.syy_bc = W.sxy_bc = W.syx_bc = 0.0;
    }
    {
        // wall histories: rows that are all +0.0 and stay so are not rewritten
        const bool had = S.wnz[f] != 0;
        const bool now = (__double_as_longlong(W.deltat_bc1) | __double_as_longlong(W.deltat_bc2) |
                          __double_as_longlong(W.theta_bc1) | __double_as_longlong(W.theta_bc2)) != 0;
        if (had || now) {
            S.wbc[0 * SS + f] = W.deltat_bc1; S.wbc[1 * SS + f] = W.deltat_bc2;
            S.wbc[2 * SS + f] = W.theta_bc1; S.wbc[3 * SS + f] = W.theta_bc2;
        }
        if (had != now) S.wnz[f] = now;
    }
    double a[NACC];
#pragma unroll
    for (int c = 0; c < NACC; ++c) {
        const bool bond_row = c == A_FBX || c == A_FBY || c == A_MB || c == A_TAB || c == A_PB;
        const bool shel_row = c == A_SHA || c == A_SHW;
        const bool stress_row = c >= A_SXX;
        const bool valid = stress_row ? REC
                         : shel_row || (bond_row ? (cl & CLS_BOND) : (cl & CLS_CONTACT));
        a[c] = valid ? S.acc[c * SS + k] : 0.0;
    }
    // bond terms: rows the fix-up wrote (CLS_BOND), else gathered here
    if (!(cl & CLS_BOND) && bg) gather_bonds(S, k, ni, no, a);
    // shelter minima (reset_shelter + thread reduction); with shelter off
    // the reference's minima stay at their zero initial value
    double sha = P.shelter ? a[A_SHA] : 0.0;
    double shw = P.shelter ? a[A_SHW] : 0.0;
    const double2 w = wind_at(S, k);
    Forcing Fo;
    forcing<MIXED>(P, F, w.x, w.y, P.uw, P.vw, sha, shw, Fo);
    double tfx = a[A_FCX] + a[A_FBX] + Fo.fax + Fo.fwx + Fo.fcorx + W.fx_bc;
    double tfy = a[A_FCY] + a[A_FBY] + Fo.fay + Fo.fwy + Fo.fcory + W.fy_bc;
    double mt = a[A_MC] + a[A_MB] + Fo.ma + Fo.mw + W.m_bc;
    double u = F.u + (tfx / F.mass) * P.dt;
    double v = F.v + (tfy / F.mass) * P.dt;
    double om = F.omega + (mt / F.inertia) * P.dt;
    FloeRec out;
    out.x = F.x + u * P.dt;
    out.y = F.y + v * P.dt;
    out.u = u; out.v = v; out.omega = om;
    out.r = F.r; out.h = F.h; out.mass = F.mass; out.inertia = F.inertia; out.hfa = F.hfa; out.hfw = F.hfw;
    out.theta = S.kin[k].theta + om * P.dt;
    {
        // next step builds the supersets anew if this floe has moved more
        // than 0.45 skin since the last build (a pair not in a superset was
        // more than its query radius + skin apart then, so it stays outside
        // the query radius while both floes stay within half the skin);
        // beyond 0.3 skin the next step with the build kernels builds early
        const double2 p0 = S.x0[k];
        const double ex = out.x - p0.x, ey = out.y - p0.y, e2 = ex * ex + ey * ey;
        const double lim = 0.45 * P.skin, soon = 0.3 * P.skin;
        // one atomic per wavefront (every such floe writing the one word
        // would queue the writes behind each other)
        const uint64_t md = __ballot(e2 > lim * lim), ms = __ballot(e2 > soon * soon);
        const int bits = (md ? BUILD_DUE : 0) | (ms ? BUILD_SOON : 0);
        const int leader = __ffsll((long long)(md | ms)) - 1;
        if (bits && (int)(threadIdx.x & 63) == leader) atomicOr(S.rebuild, bits);
    }
    if (REC) {
        double tp = (a[A_TAC] != 0.0 ? a[A_PC] / a[A_TAC] : 0.0)
                  + (a[A_TAB] != 0.0 ? a[A_PB] / a[A_TAB] : 0.0)
                  + (W.ta_bc != 0.0 ? W.p_bc / W.ta_bc : 0.0);
        // total stress: pairs + last wall + drag (forcing.f90 sig*_aw)
        const double sxx_aw = (Fo.fax + Fo.fwx) * F.r, syy_aw = (Fo.fay + Fo.fwy) * F.r;
        const double tsxx = a[A_SXX] + W.sxx_bc + sxx_aw;
        const double tsyy = a[A_SYY] + W.syy_bc + syy_aw;
        const double tsxy = a[A_SXY] + W.sxy_bc + 0.0;
        const double tsyx = a[A_SYX] + W.syx_bc + 0.0;
        double cols[33] = {tfx, tfy, mt, a[A_FCX], a[A_FCY], a[A_MC], a[A_FBX], a[A_FBY], a[A_MB],
                           Fo.fax, Fo.fay, Fo.ma, Fo.fwx, Fo.fwy, Fo.mw, Fo.fcorx, Fo.fcory,
                           W.fx_bc, W.fy_bc, W.m_bc, sha, shw, tp, tsxx, tsyy, tsxy, tsyx,
                           a[A_TAC], a[A_PC], a[A_TAB], a[A_PB], W.ta_bc, W.p_bc};
        for (int c = 0; c < 33; ++c) S.fout[c * SS + f] = cols[c];
    }
    // flags set by the repair are cleared (most floes have none: no write)
    if (dflags) S.dirty[f] = 0;
    // the new state, in the current order (the next step's copies unless it
    // rebuilds them from rec): position, kinematics, the wind at the new
    // position
    {
        GeoRec g;
        g.x = out.x; g.y = out.y; g.r = out.r; g.id = me.id; g.member = m;
        S.geo[k] = g;
        KinRec q;
        q.u = out.u; q.v = out.v; q.omega = out.omega; q.h = out.h; q.mass = out.mass;
        q.inertia = out.inertia; q.theta = out.theta;
        S.kin[k] = q;
        // the sheltering copy changes only with a reshape or a gridded wind
        // (the uniform wind stands until the host changes it, which
        // refreshes the copies)
        if (dirty || P.wind_grid >= 0) {
            double ua = P.ua, va = P.va;
            // at the start of the next step: model time tcur dt
            if (P.wind_grid >= 0)
                sample_wind(S.wgrids[P.wind_grid], S.wpool, out.x, out.y, out.r,
                            (double)S.tcur[m] * P.dt, ua, va);
            ShelRec h;
            h.hfa = out.hfa; h.hfw = out.hfw; h.wx = ua; h.wy = va;
            S.shel[k] = h;
        }
    }
    if (f == S.members[m].first) {   // once per member
        S.rebuild[2] = 1;                // the copies hold the new state
        S.members[m].tstep += 1;
        if (!member_live(S.members[m])) S.rebuild[1] = 1;   // its floes leave the next step
    }
}

// per-member totals of the last step: candidates, contacts (evf flags) and
// current bonds; diagnostics only
__global__ void k_totals(DS S, int64_t* out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t c = 0, ct = 0, bd = 0;
    if (k < S.N && S.geo[k].member >= 0) {
        int no = S.ocnt[k];
        for (int s = 0; s < no; ++s) {
            uint32_t fl = S.evf[(int64_t)s * S.S + k];
            if (fl & EV_REMOVED) continue;
            ++c;
            if (fl & EV_CONTACT) ++ct;
        }
        int f = S.cval[k];
        int nb = S.bcnt_old[f];
        for (int q = 0; q < nb; ++q) {
            const BondRec& b = S.bond_old[(int64_t)q * S.S + f];
            if (b.j >= 0 && b.bond == 1) ++bd;
        }
    }
    // wavefront reduction, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_down(c, off);
        ct += __shfl_down(ct, off);
        bd += __shfl_down(bd, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd((unsigned long long*)&out[0], (unsigned long long)c);
        atomicAdd((unsigned long long*)&out[1], (unsigned long long)ct);
        atomicAdd((unsigned long long*)&out[2], (unsigned long long)bd);
    }
}

// explicit instantiations used by the host launcher
template __global__ void k_sup<false>(DS);
template __global__ void k_sup<true>(DS);
template __global__ void k_sup_dense<false>(DS);
template __global__ void k_sup_dense<true>(DS);
template __global__ void k_cand_dense<false, false>(DS);
template __global__ void k_cand_dense<false, true>(DS);
template __global__ void k_cand_dense<true, false>(DS);
template __global__ void k_cand_dense<true, true>(DS);
template __global__ void k_cand_dense_over<false>(DS);
template __global__ void k_cand_dense_over<true>(DS);
template __global__ void k_cand<false>(DS);
template __global__ void k_cand<true>(DS);
template __global__ void k_contact<false>(DS);
template __global__ void k_contact<true>(DS);
template __global__ void k_bond<false>(DS);
template __global__ void k_bond<true>(DS);
template __global__ void k_walls<false>(DS);
template __global__ void k_walls<true>(DS);
template __global__ void k_repair<false>(DS, int);
template __global__ void k_repair<true>(DS, int);
template __global__ void k_fixup<false>(DS);
template __global__ void k_fixup<true>(DS);
template __global__ void k_final<false, false>(DS);
template __global__ void k_final<true, false>(DS);
template __global__ void k_final<false, true>(DS);
template __global__ void k_final<true, true>(DS);

}  // namespace godar
