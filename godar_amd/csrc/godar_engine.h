// Device-state layout and kernel declarations of the GODAR step pipeline.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "godar_types.h"

namespace godar {

// Accumulator rows of the per-floe gather (stepper.f90 accumulators).
enum : int {
    A_FCX = 0, A_FCY, A_MC, A_FBX, A_FBY, A_MB, A_SHA, A_SHW, A_TAC, A_PC, A_TAB, A_PB, NACC
};

constexpr int WALL_SLOT = 0xFFFF;

// Everything the kernels touch.  Per-floe arrays have length S (the floe
// capacity); per-floe lists are slot-major, entry s of floe f at [s*S + f],
// so consecutive floes read consecutive addresses.
struct DS {
    int64_t S;                 // floe capacity = stride of slot-major arrays
    int32_t N;                 // floe slots in use [0, N)
    int32_t CO, CI, HC, BC;    // outer / inner / history / bond list capacities
    int32_t SEEDK, HEAPK, TABK;  // repair pool entries per floe slot
    int32_t record_forces;
    int32_t ncells;            // cell slots (sentinel key = ncells)
    // floe state
    double *x, *y, *u, *v, *theta, *omega, *r, *h, *mass, *inertia, *hfa, *hfw;
    double* wbc;               // [4][S] deltat_bc1, deltat_bc2, theta_bc1, theta_bc2
    double *uai, *vai;         // wind at each floe this step
    int32_t* member_of;        // [S], -1 = free slot
    Member* members;
    const DevParams* params;
    const WindGridDesc* wgrids;
    const double* wpool;
    // cell list
    uint32_t *ckey, *cval;
    int32_t *cstart, *cend;
    double *sx, *sy, *sr;
    int32_t* sidx;
    // candidate lists
    int32_t *ocnt, *oj;
    double* od;
    int32_t *icnt, *ik;
    uint32_t* evf;
    int32_t* ov;
    // pair state, double-buffered
    HistRec *hist_old, *hist_new;
    int32_t *hcnt_old, *hcnt_new;
    BondRec *bond_old, *bond_new;
    int32_t *bcnt_old, *bcnt_new;
    // speculative gather
    double* acc;               // [NACC][S]
    WallOut* wout;             // [S]
    // ordered repair
    uint64_t* seeds;           // SEEDK per floe slot, region of member m starts at SEEDK*first
    int32_t* nseeds;           // [members]
    uint64_t* heap;            // HEAPK per floe slot
    EventOut* etab;            // TABK per floe slot
    int32_t* netab;
    WallOut* wtab;             // one per floe slot
    int32_t* nwtab;
    int32_t* ovw;              // [S]
    double* cur;               // [6][S] r h mass inertia hfa hfw during repair
    uint8_t *dirty, *affected; // [S]
    int64_t* stats;            // [members][4]
    double* fout;              // [33][S]
    uint32_t* err;
};

}  // namespace godar
