// hsolve_internal.h — device-side layout shared by the host engine and the
// HIP kernels.
//
// Every cell owns one contiguous region of the engine's device arena.  All
// offsets below are byte offsets into that arena (so the arena can move).
// Per-compartment arrays are indexed by the Hines index i (0..n-1).  Channel
// and gate-state arrays are "slot-major": channel slot k of compartment i sits
// at [k*n + i], gate-state slot j at [j*n + i], Ca-pool slot q at [q*n + i],
// so a wavefront walking consecutive compartments issues coalesced loads.
#pragma once
#include <stdint.h>

#include "../../include/hsolve_amd.h"

namespace hsg {

constexpr int kMaxSlots = 8;      // channel slots per compartment
constexpr int kMaxGateSlots = 24; // gate-state slots per compartment
constexpr int kMaxPoolSlots = 4;  // Ca-pool slots per compartment

// One channel type: everything about an HHChannel that does not vary per
// instance (hsolve/HSolveStruct.h:53-108 ChannelStruct, plus the gate columns
// and Ca wiring of hsolve/HSolveActiveSetup.cpp:235-526).
struct ChanType {
    double Xpower, Ypower, Zpower;
    int32_t instant;        // INSTANT_X=1, INSTANT_Y=2, INSTANT_Z=4
    int32_t nGates;         // state slots used (gates with power > 0)
    int32_t colX, colY, colZ;   // column (2*species) in the cell's table set
    int32_t caDepLocal;     // Z gate reads this local pool's row (-1: none)
    int32_t caTgtLocal;     // IkOut feeds this local pool (-1: none)
    int32_t tableSet;
    int32_t pad;
};

// One lookup table (hsolve/RateLookup.h LookupTable).
struct TableDev {
    double  min, max, dx, invDx;
    int32_t nPts, nColumns;
    int64_t off64;          // f64 table [nPts][nColumns]
    int64_t off32;          // f32 copy (mixed precision)
};

struct TableSetDev {
    TableDev v, ca;
};

// One Interpol2D table (builtins/Interpol2D.cpp), shared engine-wide.
struct Table2DDev {
    double  xmin, xmax, ymin, ymax, invDx, invDy;
    int32_t nx, ny;             // xdivs + 1, ydivs + 1
    int64_t off;                // f64 [nx * ny], x-major; a pair record: (A, B) as f64 pairs
    int64_t off32;              // a pair record: (A, B) as f32 pairs; a single table: 0
};

// What HHChannel2D objects of one kind share (biophysics/HHChannel2D.cpp):
// gate powers, tables, table axes, instant flags.  One record per distinct
// kind in the cell, read from cache by every instance.
struct Chan2DType {
    double  pw[3];
    int32_t tabA[3], tabB[3];   // engine-wide Table2DDev ids (gate with power > 0)
    int32_t dep[6];             // table axes of gate g: dep[2g], dep[2g+1] (0 Vm, 1 conc1, 2 conc2, -1)
    int32_t instant;
    int32_t pair[3];            // Table2DDev id of gate g's (A, B) pair record where the two
                                // tables share one grid, else -1
};

// One HHChannel2D run beside the solver: what varies per object.
struct Chan2DDev {
    double  Gbar, Ek, mod;
    int32_t type;               // Chan2DType of the cell
    int32_t compt;              // Hines index
    int32_t conc1, conc2;       // slot-major pool offsets of the concentrations, -1
};

// A free-running PulseGen feeding a compartment's injection
// (device/PulseGen.cpp:359-433).
struct PulseDev {
    int32_t compt;              // Hines index, -1
    int32_t entry;              // its injection entry (oInjIdx), -1
    int32_t n;                  // pulses (1..HSG_MAX_PULSES)
    int32_t trigMode;           // 0 free running, 1 external trigger, 2 external gate
    double  base, out;          // baseLevel_, output_ (state)
    double  delay[HSG_MAX_PULSES], width[HSG_MAX_PULSES], level[HSG_MAX_PULSES];
    int32_t src;                // trigMode 1-2: the free-running PulseGen driving input_, -1
    int32_t srcBefore;          // 1: src comes first in tick order (this tick's output), 0: after
    double  input, prevInput, trigTime;   // input_ (host-set where src < 0), prevInput_, trigTime_
};

// A VClamp (biophysics/VClamp.cpp:287-347): constants set at reinit, then
// the command filter, error history and current, per tick.
struct VClampDev {
    int32_t compt;              // Hines index
    int32_t entry;              // its injection entry
    int32_t mode;
    int32_t pad;
    double  Kp, expt, tauByDt, dtByTi, tdByDt;
    double  cmdIn, oldCmdIn, command, e1, e2, v1, current;
};

// A Leakage object's per-tick send: (G, G * Ek) with G = Gbar * modulation
// (biophysics/Leakage.cpp:90-95; HSolve::addGkEk).
struct LeakDev {
    double G, GE;
};

// A SynChan (biophysics/SynChan.cpp:200-227) run beside the solver.
struct SynDev {
    int32_t compt;              // Hines index
    int32_t pad;
    double  Ek, mod;
    double  xconst1, xconst2, yconst1, yconst2, norm;
    double  X, Y, Gk;           // state; Gk of the last tick
};

// One tick's activation of a SynChan (what its SimpleSynHandler sent,
// synapse/SimpleSynHandler.cpp:104-112), prepared by the host per call.
struct ActDev {
    int64_t step;               // 1-based tick since reinit
    double  value;
};

// One block of a large cell's elimination schedule (split mode): a set of
// subtrees of the unit tree, eliminated by one workgroup with its rows and
// descriptors in LDS.  Local rows 0..nRows-1 are the rows its units
// eliminate, nRows..nRows+nExt-1 the rows its subtrees' roots feed
// (eliminated later by the cell's top units); local term pairs 0..nT-1 stay
// inside the block, nT..nT+nTExt-1 are the pairs its roots leave for the
// top; nF fill entries; nHJ junction coefficients of its groups of four
// or five members.  The cell's top (the units above its blocks, ending
// at the root row, its only fed row) is one more such record after the
// blocks, whose nTImp term pairs (all of its pairs) are staged from HBM,
// where the blocks left theirs.  Arrays are arena offsets; descriptors use
// local rows, pairs and fills.
struct BlockDev {
    int32_t nRows, nExt, nT, nTExt, nF, nLevels, nUnits, ldsBytes, nTImp, nHJ;
    int64_t oRows;      // int32[nRows + nExt]: global row of each local row
    int64_t oTExt;      // int32[nTExt]: global term pair of each exported pair
    int64_t oTImp;      // int32[nTImp]: global term pair of each local pair (top)
    int64_t oFill;      // int32[nF]: global fill entry of each local fill
    int64_t oHJ;        // int32[nHJ]: global junction coefficient of each local one
    int64_t oLevel;     // int32[kBlockForms * nLevels + 1]: per level the start of
                        // each form (two or three members, four, five), then the end
    int64_t oDesc;      // int32[kDescInts * nUnits]: local descriptors, level order
};

// Per-cell descriptor (read by every workgroup working on the cell).
// One channel slot of a uniform cell for the fast channel pass, packed in
// three words (read with scalar loads):
//   w0: column of X | column of Y << 16
//   w1: column of Z | gate-state slot of X << 16 | of Y << 21 | of Z << 26 |
//       live << 31   (column and slot 0 when the gate is absent)
//   w2: power code of X | of Y << 3 | of Z << 6 (0: no gate, 1-4: integer
//       power) | instant bits << 9 | (Z's local pool + 1) << 12
// Only cells whose powers are all integers 1-4 (fastOk) take the fast pass.
constexpr int kFastWords = 3;

struct CellDev {
    int32_t n;              // compartments
    int32_t nChanSlots;     // channel slots per compartment (CHM)
    int32_t nGateSlots;     // gate-state slots per compartment (GM)
    int32_t nPoolSlots;     // Ca-pool slots per compartment (PM)
    int32_t uniform;        // 1: every compartment has the channel types slotType[]
    int32_t slotType[kMaxSlots];
    int32_t fastOk;                // uniform, integer powers 1-4: fast channel pass
    int32_t fast[kMaxSlots][kFastWords];
    int32_t hasModulation;  // 1: multiply by per-channel modulation
    int32_t hasPools;
    int32_t caAdvance;
    int32_t nLevels;
    int32_t nUnits;
    // levels [chainLevel, nLevels) hold one unit each (the trunk next to the
    // root): the fused kernels run them as one chain on one thread, their
    // descriptors staged in LDS after the level table (chainLevel == nLevels:
    // no chain)
    int32_t chainLevel;
    int32_t nHJ;            // doubles of junction coefficients
    int32_t nTerm;          // deferred elimination terms (pairs)
    int32_t nFill;          // fill slots (one per three-member group)
    int32_t nInj;           // injection entries
    int32_t nSpike;         // spike generators
    int32_t nProbe;         // Vm probes
    int32_t workInLds;      // 1: elimination work arrays live in LDS
    int32_t live;
    int32_t nGroups;
    int32_t tableSet;       // the cell's lookup-table set
    int64_t probeCap;       // samples per probe
    int64_t probeStride;    // a sample every probeStride steps (>= 1)
    // compartment arrays (f64 unless noted)
    int64_t oV, oInitVm, oCmByDt, oEmByRm, oDbase, oOff;
    int64_t oDefer;         // int32[n]: first deferred-term pair of node, -1
    // channels: slot-major [k*n + i]
    int64_t oChType;        // int32 (non-uniform cells)
    int64_t oGbar, oEk;     // GT
    int64_t oMod;           // f64
    int64_t oExtCa;         // f64
    int64_t oState;         // GT, [j*n + i]
    // pools: slot-major [q*n + i]
    int64_t oPoolC, oPoolCa, oPoolBasal, oPoolF1, oPoolF2, oPoolCeil, oPoolFloor, oPoolICa;
    int64_t oPoolCount;     // int32[n]
    // Hines schedule
    int64_t oLevelStart;    // int32[nLevels+1] into units
    int64_t oUnits;         // int32: >=0 linear link i->i+1, <0 group -(u+1)
    int64_t oGrpMem;        // int32 member lists
    int64_t oGrpInfo;       // int32[4*nGroups]: memStart, size, hjOff, deferLast
    int64_t oHJ0;           // f64[nHJ] initial junction coefficients (HJCopy_)
    int64_t oDesc;          // int32[8*nUnits] unit descriptors, level order
    int32_t slowHJ;         // 1: some group uses the slow path (HJ scratch in oWork)
    int32_t ldsBytes;       // LDS this cell needs (work + level table)
    int32_t uni2d;          // 1: the objects beside the solver are one HHChannel2D per
                            // compartment (instance i on compartment i), all of one kind
    int32_t hjInLds;        // 1 (with workInLds): the groups' junction coefficients are
                            // staged into LDS after the work arrays each tick and
                            // worked on in place there
    // work (global fallback when not in LDS): D[n], B[n], HJ[nHJ], term[2*nTerm]
    int64_t oWork;
    // injections: int32 idx[nInj] + f64 basal[nInj] + f64 varying[nInj]
    int64_t oInjIdx, oInjBasal, oInjVar;
    // spike generators: int32 compt, int32 edge, f64 thr, f64 refr, f64 last, int32 fired
    int64_t oSpkCompt, oSpkEdge, oSpkThr, oSpkRefr, oSpkLast, oSpkFired;
    // probes: int64 arena offset of the sampled double [nProbe], f64
    // trace[probeCap*nProbe], int64 base: the step count at which the
    // current trace started (written by the host when it reads the trace)
    int64_t oProbeOff, oProbeTrace, oProbeBase;
    // per-call external currents: int32 idx + f64 Gk + f64 GkEk (nExt entries)
    int64_t oExtIdx, oExtG, oExtGE;
    int32_t nExt;
    int32_t cellId;
    int64_t steps;          // steps done since reinit (device copy kept in oSteps)
    int64_t oSteps;         // int64 counter
    // message-passing solve (split mode, msg == 1)
    int32_t msg;
    int32_t nSlots;
    int32_t rootSlot;
    int32_t rootCnt;
    int64_t oMsgDesc;       // int32[kMsgDescInts * nUnits]
    int64_t oRowPos;        // int32[n]: unit-order position of each row
    int64_t oDA, oBA;       // f64[2 * nUnits + 3]: rows in unit order, the root
                            // row, two dummy rows; the eliminated rows after
                            // the forward pass
    int64_t oFillU;         // f64[nUnits + 1]: A[1][2] after member 0 (three-member units)
    int64_t oVMid;          // f64[n + 1]
    // HHChannel2D objects beside the solver
    int32_t n2d;
    int32_t values2dSent;   // 1 once the solver has sent Vm/[Ca] (reinit or step)
    int64_t o2dStart;       // int32[n + 1]: channels of compartment i
    int64_t o2dChan;        // Chan2DDev[n2d]
    int64_t o2dType;        // Chan2DType[n2dType]
    int64_t o2dState;       // f64[3 * n2d]: X, Y, Z
    int64_t o2dGk;          // f64[n2d]: Gk of the last tick
    int64_t o2dExt;         // f64[2n]: external (Gk, Gk*Ek) of the last tick per compartment
    // PulseGens
    int32_t nPulse;
    int32_t pad2;
    int64_t oPulse;         // PulseDev[nPulse]
    // objects sending (Gk, Ek) every tick: Leakage, HHChannel2D, SynChan.
    // Compartments that have any (oXFlag) take their external current from
    // o2dExt, which the external-objects kernel fills before each tick
    int32_t nXObj;          // n2d + nLeak + nSyn
    int32_t nLeak;
    int32_t nSyn;
    int32_t pad3;
    int64_t oXFlag;         // int32[n]
    int64_t oLeakStart;     // int32[n + 1]
    int64_t oLeak;          // LeakDev[nLeak]
    int64_t oSynStart;      // int32[n + 1]
    int64_t oSyn;           // SynDev[nSyn]
    int64_t oActCur;        // int32[nSyn]: next activation entry of the call
    int64_t oActEnd;        // int32[nSyn]
    int64_t oAct;           // ActDev[]: the call's activations, by synchan then step
    // VClamps
    int32_t nVClamp;
    int32_t nSetupEntries;  // injection entries at setup (the ones objects send to)
    int64_t oVClamp;        // VClampDev[nVClamp]
    // per setup entry e, the PulseGens / VClamps sending to it:
    // idx[start[e] .. start[e+1]) (desc order)
    int64_t oPulseStart, oPulseIdx, oVcStart, oVcIdx;
    int64_t oInjOf;         // int32[n]: injection entry of compartment i, -1
    // block form of the split solve (large cells): nBlocks blocks and the
    // top (BlockDev records at oBlocks, the top last)
    int32_t nBlocks;
    int32_t blockLds;       // largest LDS of its blocks
    int32_t topLds;         // LDS of its top
    int32_t pad4;
    int64_t oBlocks;
};

// Elimination unit descriptor: 12 int32 (48 bytes), read as three int4 loads.
// Units of up to three members share one code path: a linear link x -> x+1
// is a two-member unit with members (x, x+1) and coefficient HS[4x+1].
//   i0: bits 0-3 member count s, bit 4 slow path, bits 8-11 number of
//       deferred term pairs of member 1 (applied before it is pivoted)
//   i1: deferred slot for the terms of the last member (-1: apply directly)
//   i2: first deferred slot of member 1
//   i3: m0 | m1 << 16      i4: m2      i5: fill slot (s == 3)
//   i6..i11: a01, a02, a12 as doubles (the initial off-diagonal entries,
//            HJCopy_ / HS[4x+1], computed at setup as the reference does)
//   groups of more than kFastGroupMax members take one of two other forms:
//   up to kDescGroupMax members, everything in the descriptor (bit 5):
//     i1: deferred slot for the last member's terms (-1: apply directly)
//     i2: offset of the group's junction coefficients in HJ
//     i3: m0 | m1 << 16   i4: m2 | m3 << 16   i5: m4
//     i6..i8: deferred terms of members 1..3: slot | count << 24, -1: none
//   larger groups (bit 4): i1 = group index, everything else from memory
// Message-passing unit descriptor (split mode): 12 int32 (48 bytes).
//   i0: s (bits 0-3) | cnt0 << 4 | cnt1 << 6: term pairs the feeders of
//       members 0 and 1 deliver (0: no feeder, 1: a two-member feeder,
//       2: a three-member feeder)
//   i1: slot of member 0 (its feeder's terms in, its VMid out), -1: none
//   i2: slot of member 1 (three-member units), -1: none
//   i3: slot of the last member (this unit's terms out, its VMid in)
//   i4: m0 | m1 << 16 (Hines indices of the eliminated members; m1 = m0 in
//       two-member units)
//   i5: 0
//   i6..i11: a01, a02, a12
// The rows D, B of a compartment live at its unit-order position: 2u + k for
// member k of the unit u that eliminates it, 2 * nUnits for the root.
constexpr int kMsgDescInts = 12;
constexpr int kSlotDoubles = 4;    // two (tD, tB) pairs; VMid reuses the first

constexpr int kFastGroupMax = 3;
constexpr int kDescGroupMax = 5;   // groups held in their descriptor (register form)
constexpr int kAutoSplitMin = 4096;  // automatic mode: cells from this size run split
constexpr int kChannelTile = 256;  // compartments per workgroup of the channel kernel
constexpr int kPairMaxCells = 512;  // split bins of up to this many cells run as pipelined pairs
constexpr int kWideSolveMin = 4096;  // split bins with a cell from this size solve with 512 threads per cell
constexpr int kDescInts = 12;
constexpr int kMaxChain = 64;       // units of a cell's one-thread chain (their descriptors sit in LDS)
// LDS of one block (rows, exported rows, terms, fills, levels): three per CU
// (156 of 160 KiB), with room for the static LDS of the kernels beside them
constexpr int kBlockMaxLds = 52 * 1024;
constexpr int kBlockForms = 3;           // unit forms per block level: 2-3 members, 4, 5

struct SpikeRec {
    int32_t cell;
    int32_t gen;
    int64_t step;
};

struct LaunchArgs {
    char*          arena;
    const CellDev* cells;       // device array of CellDev
    const int32_t* order;       // cells to run (indices into cells)
    const ChanType* types;
    const TableSetDev* tsets;
    const Table2DDev* tabs2d;
    SpikeRec*      spikes;
    unsigned long long* spikeCount;
    int32_t        spikeCap;
    int32_t        nsteps;      // steps in this launch
    int32_t        firstOfCall; // 1: apply the call's one-shot inputs at step 0
    int32_t        ldsBytes;    // dynamic LDS per workgroup
    double         dt;          // clock tick dt
    int32_t        deferD;      // split mode: the solve kernels stop at VMid; the
                                // update and outputs kernels finish the tick
    int32_t        fp32;        // mixed form: FP32 channel tables and currents
    int32_t        ext2dInRow;  // split mode, cells with one HHChannel2D per
                                // compartment (uni2d): the channel kernel
                                // advances it in place of the external kernel
};

}  // namespace hsg
