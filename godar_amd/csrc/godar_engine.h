// Device-state layout and kernel declarations of the GODAR step pipeline.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "godar_types.h"

namespace godar {

// Accumulator rows of the per-floe gather (stepper.f90 accumulators).
enum : int {
    A_FCX = 0, A_FCY, A_MC, A_FBX, A_FBY, A_MB, A_SHA, A_SHW, A_TAC, A_PC, A_TAB, A_PB,
    A_SXX, A_SYY, A_SXY, A_SYX,   // stress diagnostics (stepper.f90:238-271)
    NACC
};

constexpr int WALL_SLOT = 0xFFFF;
// cls bits: which accumulator rows of a floe hold gathered values
// (evf words keep EV_* flags in the low 16 bits and, for bonded events,
// 1 + the owner's bond record slot in the high 16 bits)
constexpr uint8_t CLS_CONTACT = 1, CLS_BOND = 2;
constexpr int FLOE_BLOCK = 256;   // upper bound of the per-floe kernels' block size
constexpr int XCD_COUNT = 8;      // accelerator complex dies of an MI355X
constexpr int NQUEUE = 5;         // work queues: contact, bond, wall, dense search, fixup
constexpr int QN_STRIDE = 64;     // ints between queue counters (256 B)
constexpr int SCAN_TILE = 1024;   // positions per tile of the superset offset scan
constexpr int QUEUE_GRID_WAVES = 4096;   // wavefronts launched for a queue kernel at most
constexpr int FIXUP_GRID_WAVES = 2048;   // wavefronts of the fix-up (about one resident round)

// Per-floe record (96 bytes) by storage slot.  Between builds of the cell
// order the floe state lives in the ordered copies below, which the step
// reads and writes; the records are brought up to date from them (k_spill)
// before anything reads or writes them: on steps that build or refresh the
// copies, and before the host reads floes, adds floes or removes a member.
struct FloeRec {
    double x, y, u, v, omega;
    double r, h, mass, inertia, hfa, hfw;
    double theta;
};
// The floes are held in cell order ("pos"), so that a floe's neighbours are
// close in memory, and split by what the kernels read:
//   GeoRec  - neighbour search and every pair test (32 B)
//   ShelRec - sheltering: draft coefficients and the wind at the floe (32 B)
//   KinRec  - contact, bond, wall and forcing terms, integration (56 B)
struct GeoRec {
    double x, y, r;
    int32_t id;      // floe identity (member first + the caller's index)
    int32_t member;  // -1: free slot
};
struct ShelRec { double hfa, hfw, wx, wy; };
struct KinRec { double u, v, omega, h, mass, inertia, theta; };

// Per-position step flags (wflag): the floe touched a wall this step (wout
// holds its terms); the repair recomputed its walls (wtab entry in ovw); the
// repair touched the floe, so k_fixup re-gathers it.
constexpr uint32_t WF_WALL = 1, WF_REPAIR_WALL = 2, WF_FIXUP = 4;
// dirty bits (by storage slot, during the repair): the floe was reshaped (cur
// rows 0-5 hold its properties); its iteration's query radius is kept (cur
// row 6).  k_final clears them.
constexpr uint32_t DIRTY_SHAPE = 1, DIRTY_RQ = 2;

// Everything the kernels touch.  "id" arrays are indexed by a floe's storage
// slot; "pos" arrays by the floe's position in this step's cell order.  Per-floe lists
// are slot-major, entry s of row r at [s*S + r].
struct DS {
    int64_t S;                 // floe capacity = stride of slot-major arrays
    int32_t N;                 // floe slots in use [0, N)
    int32_t CO, CI, HC, BC;    // outer / inner / history / bond list capacities
    int32_t SEEDK, HEAPK, TABK;  // repair pool entries per floe slot
    int32_t record_forces;
    int32_t ncells;            // cell slots (sentinel key = ncells)
    int32_t qseg;              // floe positions per XCD share of the candidate search
    int32_t repair_lanes;      // events the repair evaluates at once (1: one at a time)
    int32_t repair_cap;        // pending repair keys held in LDS
    int32_t pad_r_;
    // floe state by storage slot (current where rebuild[2] is clear, else
    // in the ordered copies)
    FloeRec* rec;
    double* wbc;               // [4][S] deltat_bc1, deltat_bc2, theta_bc1, theta_bc2
    int32_t* member_of;        // [S], -1 = free slot
    // A floe's identity (member first + its index in the caller's arrays)
    // orders the reference's sequential loop and names it in pair records;
    // its storage slot indexes the persistent arrays.  Members store their
    // floes in the cell order of their first step, so that a step's
    // cell-ordered sweep reads and writes storage nearly sequentially.
    int32_t* gid_of;           // [S] storage slot -> identity
    int32_t* sid_of;           // [S] identity -> storage slot
    Member* members;
    const DevParams* params;
    const WindGridDesc* wgrids;
    const double* wpool;
    // cell list
    uint32_t *ckey, *cval;     // sort keys / floe ids, sorted (cval[pos] = storage slot)
    uint32_t *bkey, *bval;     // the sort's input, written by k_begin (the sort's scratch)
    int2* crange;              // [ncells+1] (start, end) in cell order
    GeoRec* geo;               // [S] pos
    ShelRec* shel;             // [S] pos (wind at the floe this step)
    KinRec* kin;               // [S] pos
    int32_t* pos_of;           // [S] storage slot -> pos
    int32_t* pos_id;           // [S] identity -> pos (floes of live members)
    // candidate lists (pos rows; entries are partner positions)
    int32_t* ocnt;
    int64_t* oj;
    double* od;
    int32_t* icnt;
    int64_t* ik;
    uint32_t* evf;
    int32_t* ov;               // repair table entry, valid where evf has EV_TABLE
    double *ssc, *ssb;         // [CO][S] pos: the pair's |(fcn,fct)| and |(fbn,fbt)|
                               // (written only when forces are recorded)
    // pair state, double-buffered (id rows; partner stored as id)
    HistRec *hist_old, *hist_new;
    int32_t *hcnt_old, *hcnt_new;
    BondRec *bond_old, *bond_new;
    int32_t *bcnt_old, *bcnt_new;
    BondTerms* bterm;          // [BC][S] (pos): this step's terms of owner record q
    uint8_t* bgat;             // [S] pos: 1 if the floe has a bonded event this step
    int32_t* ibond;            // [CI][S] pos: 1 + owner's record slot of the inner pair's bond
    // speculative gather (pos)
    double* acc;               // [NACC][S]
    WallOut* wout;             // [S], valid where WF_WALL is set
    uint8_t* wflag;            // [S rounded up to 4] WF_* bits
    int32_t* ovw;              // [S] repair wall-table entry, valid where WF_REPAIR_WALL is set
    uint8_t* cls;              // [S] CLS_* bits
    // Candidate supersets: every floe within its query radius plus the
    // member's skin at the last build, (id << 32 | pos), floe k's at
    // sup[soff[k] .. soff[k] + scnt[k]).  They are rebuilt (cell binning,
    // sort, search) only on steps that start with *rebuild set: after a
    // change of members, or once some floe has moved more than 0.45 skin from
    // where it was at the build (x0); in between, each step's candidate lists
    // are the superset entries within the query radius, the same lists a full
    // search gives (radii only shrink).
    int64_t* sup;              // [sup_cap]
    int64_t sup_cap;
    int32_t* scnt;             // [S] by pos
    int32_t* sout;             // [S] by pos, dense members: the superset holds its outer
                               // entries (partner id above the floe's) first, by (distance
                               // at the build, id), then the inner ones by id; sout is the
                               // outer count (-1: left in scan order, too long to order)
    int64_t* soff;             // [S] by pos
    int64_t* sblk;             // [S / SCAN_TILE + 2] scan of per-tile totals
    double2* x0;               // [S] by pos (positions stand between builds)
    int32_t* rebuild;          // [3] 0: build the supersets; 1: refresh the ordered copies;
                               //     2: the copies are newer than rec
    // Work queues (contact, bond, wall, dense search, fixup), each in XCD_COUNT
    // segments: the candidate search's blocks on XCD x cover positions
    // [x qseg, (x+1) qseg) and queue them into segment x, at
    // qlist[c S + x qseg + i], counted by their own counter (one per
    // QN_STRIDE ints), so that the counters' atomics spread over many lines.
    int32_t* qlist;            // [NQUEUE][S]
    int32_t* qn;               // [NQUEUE][XCD_COUNT] x QN_STRIDE
    // ordered repair (id)
    uint64_t* seeds;           // SEEDK per floe slot, member region starts at SEEDK*first
    int32_t* nseeds;           // [members]
    uint64_t* heap;            // HEAPK per floe slot
    EventOut* etab;            // TABK per floe slot
    int32_t* netab;
    WallOut* wtab;             // one per floe slot
    int32_t* nwtab;
    double* cur;               // [7][S] r h mass inertia hfa hfw, query radius during repair
    uint8_t* dirty;            // [S rounded up to 4] DIRTY_* bits
    int32_t* comp;             // [S] by pos: contact component (union-find parent)
    uint8_t* wnz;              // [S] by storage slot: 0 = the floe's wall history rows (wbc)
                               //     are all +0.0 (k_final then skips rewriting them)
    int64_t* stats;            // [members][4]
    int32_t* tcur;             // [members] number of the step in progress (set at its start)
    double* fout;              // [33][S] (id)
    uint32_t* err;
    int64_t* steps_done;       // steps completed without error (whole context)
};

}  // namespace godar
