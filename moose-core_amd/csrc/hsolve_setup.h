// hsolve_setup.h — host-side model reading for one cell: the work of
// HSolve::setup (hsolve/HSolve.cpp:286-293) turned into the flat, Hines-ordered
// arrays and the parallel elimination schedule that the GPU kernels run.
#pragma once
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/hsolve_amd.h"

namespace hsg {

// One lookup table as HSolveActive::createLookupTables builds it
// (hsolve/HSolveActiveSetup.cpp:339-526, hsolve/RateLookup.cpp:15-55).
struct HostTable {
    double min = 0, max = 0, dx = 0;
    uint32_t nDivs = 0, nSpecies = 0;
    std::vector<double> table;   // [nDivs+2][2*nSpecies]
    uint32_t nPts() const { return nDivs + 2; }
    uint32_t nColumns() const { return 2 * nSpecies; }
};

struct HostTableSet {
    HostTable v, ca;
    std::string key;
};

// One Interpol2D (builtins/Interpol2D.cpp:173-191 for invDx/invDy).
struct HostTable2D {
    uint32_t nx = 0, ny = 0;          // xdivs + 1, ydivs + 1
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0, invDx = 1, invDy = 1;
    std::vector<double> table;        // [ix * ny + iy]
    uint64_t hash = 0;
};

// One HHChannel2D (device record; see Chan2DDev).
struct HostChan2D {
    int compt;                        // Hines index
    int desc;                         // index in hsg_cell_desc.chans2d
    int instant;
    int tabA[3], tabB[3];             // local 2-D table indices, -1
    int dep[6];
    int conc1, conc2;                 // slot-major pool offsets (q*n + i), -1
    double Gbar, Ek, mod, pw[3], st[3];
};

struct HostPulse {
    int compt;                        // Hines index, -1 (drives another PulseGen's input only)
    int entry;                        // injection entry (Hines order of injected compartments), -1
    int n;
    int trigMode;                     // 0 free running, 1 external trigger, 2 external gate
    int src;                          // trigMode 1-2: the PulseGen driving the input, -1
    double base;
    double delay[HSG_MAX_PULSES], width[HSG_MAX_PULSES], level[HSG_MAX_PULSES];
};

// Whether PulseGen p can drive another's input on the device: free running
// with a positive period, and every phase in [0, period) inside one of its
// pulses or the gaps before them, so that device/PulseGen.cpp:408-421 sets
// the output on every tick and the output is a function of the time alone.
bool pulse_source_ok(const HostPulse& p);

// The output of such a source at time t (device/PulseGen.cpp:359-421 in
// FREE_RUN mode), for a host read of one that feeds no compartment.
double pulse_free_output(const HostPulse& p, double t);

// A VClamp (biophysics/VClamp.cpp) sending to its compartment's injectMsg.
struct HostVClamp {
    int compt;                        // Hines index
    int entry;                        // injection entry
    int mode;
    double tau, ti, td, gain;
};

// A Leakage object's per-tick send (biophysics/Leakage.cpp:90-95).
struct HostLeak {
    int compt;                        // Hines index
    int desc;                         // index in hsg_cell_desc.leaks
    double Gbar, Ek, mod;
};

// A SynChan with its handler (biophysics/SynChan.cpp,
// synapse/SimpleSynHandler.cpp).
struct HostSyn {
    int compt;                        // Hines index
    int desc;                         // index in hsg_cell_desc.synchans
    int firstSyn, nSyn;               // its synapses (desc order)
    double Gbar, Ek, tau1, tau2, mod;
};

struct HostChanType {
    double Xpower, Ypower, Zpower;
    int instant, nGates, colX, colY, colZ, caDepLocal, caTgtLocal;
    std::string tableKey;
};

// Elimination unit: linear link (node) or junction group.
struct Unit {
    int kind;        // 0 linear, 1 group
    int node;        // linear: eliminated node x (x -> x+1)
    int group;       // group index
    int level;
};

struct Group {
    std::vector<int> mem;   // sorted Hines indices
    int hjOff;
    int deferLast;          // term offset for the last member, -1
};

struct CellSetup {
    int n = 0;
    double dt = 0;                    // setup dt (HinesMatrix::dt_)
    int caAdvance = 1;
    std::vector<int> hinesToDesc, descToHines;
    // tree_ and compartment_ (hsolve/HSolvePassive.cpp:100-205)
    std::vector<double> Ra, Rm, Cm, Em, initVm, V, CmByDt, EmByRm;
    std::vector<double> injBasal;     // per Hines index
    std::vector<char> hasInject;
    // Hines matrix (hsolve/HinesMatrix.cpp:136-222)
    std::vector<double> Dbase, Off, HJ0;
    std::vector<Group> groups;
    std::vector<int> units;           // encoded, level-ordered
    std::vector<int> levelStart;
    // unit descriptors, 8 int32 per unit in level order (see UnitDesc in
    // hsolve_internal.h)
    std::vector<int32_t> udesc;
    bool anySlowGroup = false;
    // message-passing form of the schedule (every unit on the fast path):
    // per-unit descriptors (kMsgDescInts each, level order), the position of
    // each compartment's matrix row in unit order, and the LDS message slots
    bool msgOk = false;
    std::vector<int32_t> mdesc;
    std::vector<int32_t> rowPos;      // per Hines index
    int nSlots = 0;
    int rootSlot = -1;
    int rootCnt = 0;
    std::vector<int> defer;           // 2n: (offset, count)
    int nTerm = 0;
    int nFill = 0;
    int maxGroup = 0;
    // channels, slot-major
    int chanSlots = 0, gateSlots = 0, poolSlots = 0;
    bool uniform = true;
    std::vector<int> slotTypeUniform;        // per slot when uniform
    std::vector<int> chType;                 // [k*n + i] engine-wide type id, -1
    std::vector<double> Gbar, Ek, mod, state;  // slot-major
    bool hasModulation = false;
    // desc channel -> (compartment Hines index, slot k, first state slot j)
    std::vector<int> chanCompt, chanSlot, chanState;
    std::vector<double> chanXp, chanYp, chanZp;
    // pools
    std::vector<int> poolCount;              // per compartment
    std::vector<int> poolCompt, poolSlot;    // per desc pool (-1 unused)
    std::vector<double> pC, pCa, pBasal, pF1, pF2, pCeil, pFloor;  // slot-major
    // spike generators (Hines order of their compartments)
    std::vector<int> spkCompt, spkDesc;
    std::vector<int> spkEdge;
    std::vector<double> spkThr, spkRefr;
    // types used (engine-wide ids filled by the engine)
    // HHChannel2D objects beside the solver, sorted by compartment (desc
    // order within one), with the compartment -> channels index (n + 1)
    std::vector<HostChan2D> ch2d;
    std::vector<HostPulse> pulses;    // PulseGens on injectMsg, desc order
    std::vector<HostVClamp> vclamps;  // VClamps on injectMsg, desc order
    // the injection entries at setup, and per entry the PulseGens and
    // VClamps sending to it (indices, desc order within an entry)
    int nSetupEntries = 0;
    std::vector<int32_t> pulseStart, pulseIdx, vcStart, vcIdx;
    // Leakage sends and SynChans, sorted by compartment (desc order within
    // one), with their compartment -> objects indexes (n + 1)
    std::vector<HostLeak> leaks;
    std::vector<int32_t> leakStart;
    std::vector<int> leakOfDesc;
    std::vector<HostSyn> syn;
    std::vector<int32_t> synStart;
    std::vector<int> synOfDesc;
    std::vector<double> synW, synDelay;      // synapses, desc order
    std::vector<int32_t> ch2dStart;
    std::vector<int> ch2dOfDesc;             // desc index -> position in ch2d
    std::vector<HostTable2D> tables2d;
    // block form of the elimination schedule (large cells, split mode):
    // blocks of subtrees of the unit tree with local rows/terms/fills, and
    // the remaining (top) units in the ordinary format
    struct Block {
        std::vector<int32_t> rows;     // global rows: eliminated (nRows), then fed (nExt)
        int nRows = 0, nExt = 0, nT = 0;
        std::vector<int32_t> tExt;     // global term pair of each exported pair
        std::vector<int32_t> tImp;     // global term pair of each local pair (the top)
        std::vector<int32_t> fills;    // global fill entry of each local fill
        std::vector<int32_t> hj;       // global junction coefficient of each local one
        std::vector<int32_t> level;    // local level table
        std::vector<int32_t> desc;     // local descriptors, level order
        int64_t ldsBytes = 0;
    };
    std::vector<Block> blocks;         // the blocks, then the top
    std::vector<HostChanType> localTypes;    // indexed by local type id
    std::vector<int> chTypeLocal;            // [k*n+i] local type id
    HostTableSet tables;
};

// Reads one cell.  Returns 0 or an HSG_E_* code with *err set.
int setup_cell(const hsg_cell_desc* d, double dt, int caAdvance, CellSetup& out, std::string& err);

}  // namespace hsg
