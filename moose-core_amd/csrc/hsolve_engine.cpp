// hsolve_engine.cpp — the C ABI (include/hsolve_amd.h) over HIP.
//
// An engine owns one HIP stream and one device arena.  Each cell lives in its
// own region of the arena (first-fit allocation, so cells can join and leave a
// running batch).  hsg_step launches one workgroup per live cell and fuses
// `steps_per_launch` ticks into each launch; optionally the steady-state
// launches are captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <set>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hsolve_amd.h"
#include "hsolve_internal.h"
#include "hsolve_setup.h"

extern "C" hipError_t hsg_launch_step(const hsg::LaunchArgs* a, int nblocks, int threads, int fp32, int slow,
                                     int fastms, int lean, hipStream_t stream);
extern "C" hipError_t hsg_launch_reinit(const hsg::LaunchArgs* a, int nblocks, int fp32,
                                       hipStream_t stream);
extern "C" hipError_t hsg_set_max_lds(int bytes);
extern "C" hipError_t hsg_launch_solve(const hsg::LaunchArgs* a, int nblocks, int threads, int fp32, int slow,
                                      hipStream_t stream);
extern "C" hipError_t hsg_launch_channel(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, int fp32,
                                        hipStream_t stream);
extern "C" hipError_t hsg_launch_external(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, hipStream_t stream);
extern "C" hipError_t hsg_launch_channel_fast(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, int fp32,
                                             int ms, int sig, hipStream_t stream);
extern "C" hipError_t hsg_launch_chan2d_reinit(const hsg::LaunchArgs* a, int nblocks, hipStream_t stream);
extern "C" hipError_t hsg_launch_update(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, int fp32, int sig,
                                       hipStream_t stream);
extern "C" hipError_t hsg_launch_outputs(const hsg::LaunchArgs* a, int nblocks, hipStream_t stream);
extern "C" hipError_t hsg_launch_msg(const hsg::LaunchArgs* a, int nblocks, int threads, int fp32,
                                    hipStream_t stream);
extern "C" hipError_t hsg_scratch_address(int device, void** ptr, long long* bytes);
extern "C" hipError_t hsg_launch_block(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, int slow, int fwd,
                                      int lds, hipStream_t stream);
extern "C" hipError_t hsg_launch_top(const hsg::LaunchArgs* a, int nblocks, int threads, int slow, int lds,
                                    hipStream_t stream);

using namespace hsg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(HSG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int64_t kAlign = 256;
constexpr int kSpikeCap = 1 << 20;

int64_t align_up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, cap) with coalescing free list.
struct Arena {
    char* base = nullptr;
    int64_t cap = 0;
    int64_t used = 0;
    std::map<int64_t, int64_t> freeList;   // offset -> size

    int64_t alloc(int64_t bytes) {
        bytes = align_up(std::max<int64_t>(bytes, kAlign));
        for (auto it = freeList.begin(); it != freeList.end(); ++it) {
            if (it->second >= bytes) {
                int64_t off = it->first;
                int64_t rest = it->second - bytes;
                freeList.erase(it);
                if (rest > 0) freeList[off + bytes] = rest;
                used += bytes;
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        bytes = align_up(std::max<int64_t>(bytes, kAlign));
        used -= bytes;
        auto it = freeList.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != freeList.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            freeList.erase(nx);
        }
        if (it != freeList.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                freeList.erase(it);
            }
        }
    }
};

struct Block {
    int64_t off = -1, bytes = 0;
};

// SynChan's integration constants and norm at step dt (setTau1, setTau2 and
// normalizeGbar, biophysics/SynChan.cpp:123-198; reinit, :229-251).
struct SynConst {
    double dt = 25.0e-6;       // SynChan's dt_ until a reinit
    double xconst1 = 0, xconst2 = 0, yconst1 = 1, yconst2 = 0, norm = 0;
};

// doubleEq (basecode/doubleEq.cpp:16-22)
bool double_eq(double x, double y) {
    double denom = std::fabs(x) + std::fabs(y);
    if (denom < 1e-12) denom = 1e-6;
    return std::fabs(x - y) / denom < 1e-6;
}

SynConst syn_constants(const HostSyn& q, double dt) {
    SynConst c;
    c.dt = dt;
    c.xconst1 = q.tau1 * (1.0 - std::exp(-dt / q.tau1));
    c.xconst2 = std::exp(-dt / q.tau1);
    if (double_eq(q.tau2, 0.0)) {
        c.yconst1 = 1.0;
        c.yconst2 = 0.0;
    } else {
        c.yconst1 = q.tau2 * (1.0 - std::exp(-dt / q.tau2));
        c.yconst2 = std::exp(-dt / q.tau2);
    }
    if (double_eq(q.tau2, 0.0)) {
        c.norm = q.Gbar;
    } else if (double_eq(q.tau1, q.tau2)) {
        c.norm = q.Gbar * std::exp(1.0) / q.tau1;
    } else {
        const double tpeak = q.tau1 * q.tau2 * std::log(q.tau1 / q.tau2) / (q.tau1 - q.tau2);
        c.norm = q.Gbar * (q.tau1 - q.tau2) /
                 (q.tau1 * q.tau2 * (std::exp(-tpeak / q.tau1) - std::exp(-tpeak / q.tau2)));
    }
    return c;
}

std::vector<LeakDev> leak_records(const CellSetup& s) {
    std::vector<LeakDev> out(s.leaks.size());
    for (size_t k = 0; k < s.leaks.size(); ++k) {
        const double G = s.leaks[k].Gbar * s.leaks[k].mod;
        out[k].G = G;
        out[k].GE = G * s.leaks[k].Ek;
    }
    return out;
}

std::vector<SynDev> syn_records(const CellSetup& s, const std::vector<SynConst>& cst,
                                const std::vector<std::array<double, 3>>& state) {
    std::vector<SynDev> out(s.syn.size());
    for (size_t k = 0; k < s.syn.size(); ++k) {
        SynDev& d = out[k];
        std::memset(&d, 0, sizeof d);
        d.compt = s.syn[k].compt;
        d.Ek = s.syn[k].Ek;
        d.mod = s.syn[k].mod;
        d.xconst1 = cst[k].xconst1;
        d.xconst2 = cst[k].xconst2;
        d.yconst1 = cst[k].yconst1;
        d.yconst2 = cst[k].yconst2;
        d.norm = cst[k].norm;
        d.X = state[k][0];
        d.Y = state[k][1];
        d.Gk = state[k][2];
    }
    return out;
}

struct HostCell {
    int id = -1;
    bool live = false;
    CellSetup s;
    CellDev dev{};
    Block region;          // main region
    Block inj, probe, extBlk; // side blocks
    double hsolveDt = 0;   // HSolve::dt_ (setup dt until reinit)
    int64_t bytes = 0;
    // host mirrors of per-call inputs (reference: inject_ map, externalCurrent_)
    std::vector<int> injIdx;           // Hines index per injection entry
    std::vector<double> injBasal, injVar;
    std::map<int, std::pair<double, double>> ext;   // Hines idx -> (sum Gk, sum Gk*Ek)
    std::vector<double> prevExtG, prevExtGE;        // for getIm
    bool prevExtSet = false;                        // some prevExt entry is nonzero
    std::map<int, double> ica;                      // desc pool -> pending iCa
    std::vector<int> typeIds;                       // local type -> engine type
    int tableSet = 0;
    std::vector<int> probeCompt;
    int64_t probeCap = 0;
    int64_t probeBase = 0;     // the device copy sits at dev.oProbeBase
    int64_t stepsDone = 0;
    bool injDirty = false;
    // Gk as of the last step for channels whose X/Y/Z, Gbar or modulation was
    // written since (the reference keeps current_[].Gk until the next step)
    std::map<int, double> gkOverride;
    // SynChans (s.syn order): integration constants at their dt, the state
    // when it was last uploaded, and the handlers' queues of (time, weight),
    // sorted by time, equal times in arrival order
    std::vector<SynConst> synConst;
    std::vector<std::array<double, 3>> synState;   // X, Y, Gk
    std::vector<std::vector<std::pair<double, double>>> synQ;
    Block act;                                     // the call's activation list
};

}  // namespace

struct hsg_engine {
    hsg_options opt{};
    int fp32 = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    Arena arena;
    std::vector<std::unique_ptr<HostCell>> cells;
    std::vector<ChanType> types;
    std::map<std::string, int> typeKey;
    std::vector<TableSetDev> tsets;
    std::map<std::string, int> tsetKey;
    std::vector<Table2DDev> tabs2d;              // Interpol2D tables, deduplicated
    std::map<uint64_t, int> tab2dKey;
    Table2DDev* dTabs2d = nullptr;
    int dTabs2dCap = 0;
    std::map<std::pair<int, int>, int> pair2dKey;  // (A, B) table ids -> pair record id
    // device-side tables of the launch
    CellDev* dCells = nullptr;
    int dCellsCap = 0;
    ChanType* dTypes = nullptr;
    int dTypesCap = 0;
    TableSetDev* dTsets = nullptr;
    int dTsetsCap = 0;
    SpikeRec* dSpikes = nullptr;
    unsigned long long* dSpikeCount = nullptr;
    int64_t spikesDropped = 0;
    // tablesDirty: the cells' CellDev records or the engine's type and table
    // lists changed and go to the device before the next launch (copies into
    // the existing buffers).  planDirty: the live set, or a property that
    // decides a cell's bin, changed, so the launch plan is rebuilt too.
    bool tablesDirty = true;
    bool planDirty = true;
    std::vector<int32_t> order;
    int ldsLimit = 65536;
    double dt = 0;                // clock tick
    bool reinitDone = false;
    // timing of the last step call
    std::vector<hipEvent_t> evStart, evStop;
    int lastLaunches = 0;
    double lastTotalMs = 0;
    // size bins of the live set, each launched on its own stream
    struct Bin {
        std::vector<int32_t> cells;
        int32_t* dOrder = nullptr;  // in the plan's device buffer
        int2* dTiles = nullptr;     // channel-kernel tiles (split mode), same buffer
        int nTiles = 0;
        bool msg = false;           // message-passing solve kernel (split mode)
        bool hasExt = false;        // cells with HHChannel2D objects: the external kernel before every tick
        bool uni2dAll = false;      // every cell has one HHChannel2D per compartment and nothing else beside
        bool split = false;         // channel kernel + solve kernel per tick
        bool slow = false;          // cells with junction groups of four or more members
        int chanMS = 0;             // channel kernel's fast pass for up to chanMS slots (0: generic)
        int64_t sumN = 0;           // compartments in the bin
        int lds = 0;
        int threads = 256;
        bool pairNext = false;      // first half of a split bin whose second half is the next bin
        int solveThreads = 256;     // threads per cell of the split solve kernel
        double roundMs = 0;         // its dispatches' duration in the last timed round
        int roundSteps = 0;         // ticks of that round
        bool blocks = false;        // split solve in block form (every cell has blocks)
        int2* dBlkTiles = nullptr;  // (cell, block) per block workgroup, in the plan's buffer
        int nBlkTiles = 0;
        int blockLds = 0;           // dynamic LDS of the block kernels
        int topLds = 0;             // dynamic LDS of the top kernel
        int key = 0;                // the bin's class (second half of a pair: + 32)
        int fastSel = -1;           // fast channel pass: -1 the engine's choice, 0 off, 1 on
        int lean = 0;               // no stimulus or external objects and no Ca pools in any cell: 1,
                                    // and every cell's channels the squid HH pair: 2
        int sig = 0;                // every cell's channel slots Na, Kdr, KA, CaL: 2 (the split channel kernel)
    };
    // Launch shapes set by the caller per bin class (hsg_set_bin_shape):
    // threads per cell workgroup (0: the engine's choice) and the fast
    // channel pass (-1: the engine's choice).
    struct BinShape {
        int threads = 0;
        int fast = -1;
    };
    std::map<int, BinShape> shapes;
    // A launch plan: the bins of a set of cells, with every bin's cell order
    // and channel tiles in one device buffer (owned by the plan; `external`
    // buffers belong to someone else and are never freed here).
    struct Plan {
        std::vector<Bin> bins;
        char* dBuf = nullptr;
        int64_t cap = 0;
        bool external = false;
    };
    Plan plan;                   // the live set's plan (hsg_step)
    // plans of cell subsets stepped by hsg_step_cells with unequal step
    // counts, keyed by the subset, most recent last; dropped with the main
    // plan whenever it is rebuilt
    std::vector<std::pair<std::vector<int32_t>, Plan>> subPlans;
    static constexpr size_t kMaxSubPlans = 8;
    bool split = false;          // channel phase as its own kernel
    static constexpr int kMaxBins = 32;
    hipStream_t binStream[kMaxBins] = {};
    hipEvent_t evFork = nullptr, evJoin[kMaxBins] = {};
    hipEvent_t evPairA[kMaxBins] = {}, evPairB[kMaxBins] = {};   // hand-offs inside a pipelined pair
    hipEvent_t evBin0[kMaxBins] = {}, evBin1[kMaxBins] = {};     // each bin's dispatches in a timed round
    // SpikeGen spikeOut -> Synapse addSpike messages between cells: (source
    // cell, spikegen desc index) -> (target cell, synchan desc index,
    // synapse index within it), in connection order
    struct SpikeTarget { int32_t cell, synchan, synapse; };
    std::map<std::pair<int32_t, int32_t>, std::vector<SpikeTarget>> conns;
    std::vector<SpikeRec> spikeBuf;   // spikes read for routing, not yet handed out
    Block actBlk;                     // the SynChans' activation list of the current call
    bool actEmpty = false;            // the list last copied held no events
    size_t actCells = 0;              // cells it covered
    // graph of the steady-state launches
    hipGraphExec_t graphExec = nullptr;
    int graphSteps = 0;
    bool graphValid = false;
    // the static launch-table area of hsg_step_cells_noalloc on this engine's
    // device, and whether the call in progress may allocate device memory
    char* scratch = nullptr;
    int64_t scratchBytes = 0;
    bool noAlloc = false;
};

namespace {

// Copies a host vector into the arena and waits, so that the caller may free
// the vector as soon as this returns.
template <typename T>
int upload(hsg_engine* e, int64_t off, const std::vector<T>& v) {
    if (v.empty()) return 0;
    HIPCHK(hipMemcpyAsync(e->arena.base + off, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice,
                          e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

// Host image of one arena region: arrays are placed into it and the region
// goes to the device in a single copy.
struct Stager {
    int64_t base;
    std::vector<char> img;
    Stager(int64_t b, int64_t bytes) : base(b), img((size_t)bytes, 0) {}
    template <typename T>
    void put(int64_t off, const std::vector<T>& v) {
        if (!v.empty()) std::memcpy(img.data() + (off - base), v.data(), v.size() * sizeof(T));
    }
    int flush(hsg_engine* e) {
        HIPCHK(hipMemcpyAsync(e->arena.base + base, img.data(), img.size(), hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        return 0;
    }
};

int grow_arena(hsg_engine* e, int64_t need) {
    if (e->opt.mem_budget > 0) return fail(HSG_E_NOMEM, "device memory budget exceeded");
    int64_t ncap = std::max<int64_t>(e->arena.cap * 2, e->arena.cap + align_up(need) + (64 << 20));
    char* nb = nullptr;
    HIPCHK(hipMalloc(&nb, ncap));
    if (e->arena.base) {
        HIPCHK(hipMemcpyAsync(nb, e->arena.base, e->arena.cap, hipMemcpyDeviceToDevice, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipFree(e->arena.base));
    }
    // the new space at the end is free
    int64_t oldCap = e->arena.cap;
    e->arena.base = nb;
    e->arena.cap = ncap;
    e->arena.freeList[oldCap] = ncap - oldCap;
    // coalesce with a trailing free block
    auto it = e->arena.freeList.find(oldCap);
    if (it != e->arena.freeList.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            e->arena.freeList.erase(it);
        }
    }
    e->graphValid = false;
    return 0;
}

int arena_alloc(hsg_engine* e, int64_t bytes, int64_t* off) {
    int64_t o = e->arena.alloc(bytes);
    if (o < 0) {
        if (e->noAlloc)
            return fail(HSG_E_NOMEM, "the engine's device arena is full and this call allocates no device memory "
                                     "(an allocating call such as hsg_step_cells grows it)");
        int rc = grow_arena(e, bytes);
        if (rc) return rc;
        o = e->arena.alloc(bytes);
        if (o < 0) return fail(HSG_E_NOMEM, "arena allocation failed");
    }
    *off = o;
    return 0;
}

template <typename T>
int grow_dev(T** p, int* cap, int need) {
    if (need <= *cap) return 0;
    int ncap = std::max(need, *cap * 2 + 16);
    if (*p) HIPCHK(hipFree(*p));
    HIPCHK(hipMalloc(p, sizeof(T) * (size_t)ncap));
    *cap = ncap;
    return 0;
}

int table_set_id(hsg_engine* e, const HostTableSet& t, int* id) {
    auto it = e->tsetKey.find(t.key);
    if (it != e->tsetKey.end()) { *id = it->second; return 0; }
    TableSetDev ts{};
    const HostTable* hs[2] = {&t.v, &t.ca};
    TableDev* ds[2] = {&ts.v, &ts.ca};
    for (int k = 0; k < 2; ++k) {
        const HostTable& h = *hs[k];
        TableDev& dv = *ds[k];
        dv.min = h.min;
        dv.max = h.max;
        dv.dx = h.dx;
        dv.invDx = 1.0 / h.dx;
        dv.nPts = (int)h.nPts();
        dv.nColumns = (int)h.nColumns();
        int64_t o64 = 0, o32 = 0;
        int64_t cnt = (int64_t)h.table.size();
        int rc = arena_alloc(e, std::max<int64_t>(cnt, 1) * 8, &o64);
        if (rc) return rc;
        rc = arena_alloc(e, std::max<int64_t>(cnt, 1) * 4, &o32);
        if (rc) return rc;
        std::vector<float> f(h.table.begin(), h.table.end());
        rc = upload(e, o64, h.table);
        if (rc) return rc;
        rc = upload(e, o32, f);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(e->stream));
        dv.off64 = o64;
        dv.off32 = o32;
    }
    *id = (int)e->tsets.size();
    e->tsets.push_back(ts);
    e->tsetKey[t.key] = *id;
    e->tablesDirty = true;
    return 0;
}

// Engine-wide id of the (A, B) pair record of two 2-D tables on one grid:
// the two tables interleaved point by point, in f64 and in f32, so that a
// lookup of both reads two neighbouring pairs per grid row.  -1 in *id when
// the grids differ.
int pair2d_id(hsg_engine* e, int idA, int idB, const HostTable2D& a, const HostTable2D& b, int* id) {
    *id = -1;
    if (a.nx != b.nx || a.ny != b.ny || a.xmin != b.xmin || a.xmax != b.xmax || a.ymin != b.ymin ||
        a.ymax != b.ymax || a.invDx != b.invDx || a.invDy != b.invDy || a.table.size() != b.table.size())
        return 0;
    const auto key = std::make_pair(idA, idB);
    auto it = e->pair2dKey.find(key);
    if (it != e->pair2dKey.end()) { *id = it->second; return 0; }
    Table2DDev d = e->tabs2d[idA];
    const size_t n = a.table.size();
    std::vector<double> p64(2 * std::max<size_t>(n, 1), 0.0);
    std::vector<float> p32(2 * std::max<size_t>(n, 1), 0.0f);
    for (size_t k = 0; k < n; ++k) {
        p64[2 * k] = a.table[k];
        p64[2 * k + 1] = b.table[k];
        p32[2 * k] = (float)a.table[k];
        p32[2 * k + 1] = (float)b.table[k];
    }
    int rc = arena_alloc(e, (int64_t)p64.size() * 8, &d.off);
    if (rc) return rc;
    if ((rc = arena_alloc(e, (int64_t)p32.size() * 4, &d.off32))) return rc;
    if ((rc = upload(e, d.off, p64)) || (rc = upload(e, d.off32, p32))) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    *id = (int)e->tabs2d.size();
    e->tabs2d.push_back(d);
    e->pair2dKey[key] = *id;
    e->tablesDirty = true;
    return 0;
}

// Engine-wide id of a 2-D table (uploaded once per content).
int table2d_id(hsg_engine* e, const HostTable2D& t, int* id) {
    auto it = e->tab2dKey.find(t.hash);
    if (it != e->tab2dKey.end()) { *id = it->second; return 0; }
    Table2DDev d{};
    d.xmin = t.xmin; d.xmax = t.xmax; d.ymin = t.ymin; d.ymax = t.ymax;
    d.invDx = t.invDx; d.invDy = t.invDy;
    d.nx = (int32_t)t.nx; d.ny = (int32_t)t.ny;
    int rc = arena_alloc(e, std::max<int64_t>((int64_t)t.table.size(), 1) * 8, &d.off);
    if (rc) return rc;
    if ((rc = upload(e, d.off, t.table))) return rc;
    *id = (int)e->tabs2d.size();
    e->tabs2d.push_back(d);
    e->tab2dKey[t.hash] = *id;
    e->tablesDirty = true;
    return 0;
}

int type_id(hsg_engine* e, const HostChanType& t, int tableSet) {
    char kb[320];
    snprintf(kb, sizeof kb, "%a|%a|%a|%d|%d|%d|%d|%d|%d|%d", t.Xpower, t.Ypower, t.Zpower, t.instant,
             t.colX, t.colY, t.colZ, t.caDepLocal, t.caTgtLocal, tableSet);
    auto it = e->typeKey.find(kb);
    if (it != e->typeKey.end()) return it->second;
    ChanType c{};
    c.Xpower = t.Xpower;
    c.Ypower = t.Ypower;
    c.Zpower = t.Zpower;
    c.instant = t.instant;
    c.nGates = t.nGates;
    c.colX = t.colX;
    c.colY = t.colY;
    c.colZ = t.colZ;
    c.caDepLocal = t.caDepLocal;
    c.caTgtLocal = t.caTgtLocal;
    c.tableSet = tableSet;
    int id = (int)e->types.size();
    e->types.push_back(c);
    e->typeKey[kb] = id;
    e->tablesDirty = true;
    return id;
}

// LDS-resident part of the elimination work: D[n], B[n], deferred terms, fills
int64_t work_lds_bytes(const CellDev& c) {
    return 8 * (2 * (int64_t)c.n + 2 * (int64_t)c.nTerm + (int64_t)c.nFill);
}
// The groups' junction coefficients when they live in LDS (after the work
// arrays, before the level table).
int64_t hj_lds_bytes(const CellDev& c) {
    return c.hjInLds ? 8 * (int64_t)c.nHJ : 0;
}
// The level table and, 16-byte aligned after it, the chain's descriptors.
int64_t level_bytes(const CellDev& c) {
    return 4 * ((int64_t)c.nLevels + 1) + 16 + 4 * kDescInts * ((int64_t)c.nLevels - c.chainLevel);
}

// Lays out and uploads one cell's main region.
void fill_fast_slots(hsg_engine* e, CellDev& d);

int place_cell(hsg_engine* e, HostCell& hc) {
    CellSetup& s = hc.s;
    const int n = s.n;
    const int gt = e->fp32 ? 4 : 8;
    const int64_t CS = (int64_t)s.chanSlots * n, GS = (int64_t)s.gateSlots * n,
                  PS = (int64_t)s.poolSlots * n;
    // sizes of every array of the region
    struct Item { int64_t* off; int64_t bytes; };
    CellDev& d = hc.dev;
    std::memset(&d, 0, sizeof d);
    int64_t nGroups = (int64_t)s.groups.size(), nMem = 0;
    for (auto& g : s.groups) nMem += (int64_t)g.mem.size();
    int64_t workB = 8 * (2 * (int64_t)n + 2 * (int64_t)s.nTerm + (int64_t)s.nFill +
                         (s.anySlowGroup ? (int64_t)s.HJ0.size() : 0));
    // the message-passing solve needs its slots (and the level table) in LDS
    const int64_t msgLds = align_up(8 * kSlotDoubles * ((int64_t)s.nSlots + 1) + 4 * (int64_t)s.levelStart.size());
    // split mode: every cell that can; automatic mode: large cells, whose
    // channel phase needs more threads than one workgroup brings
    const bool wantSplit = e->split || (e->opt.kernel_mode == HSG_KERNEL_AUTO && n >= kAutoSplitMin);
    const bool msgOn = wantSplit && s.msgOk && msgLds <= e->ldsLimit;
    // the block form of the split solve: large cells without the message
    // form whose schedule splits into blocks that fit LDS
    bool blockOn = wantSplit && !msgOn && s.blocks.size() >= 2;
    for (const auto& B : s.blocks) blockOn = blockOn && B.ldsBytes <= e->ldsLimit;
    int64_t blockBytes = 0;
    if (blockOn) {
        blockBytes = align_up((int64_t)sizeof(BlockDev) * (int64_t)s.blocks.size());
        for (const auto& B : s.blocks)
            blockBytes += align_up(4 * (int64_t)B.rows.size() + 4) + align_up(4 * (int64_t)B.tExt.size() + 4) +
                          align_up(4 * (int64_t)B.tImp.size() + 4) + align_up(4 * (int64_t)B.fills.size() + 4) +
                          align_up(4 * (int64_t)B.hj.size() + 4) +
                          align_up(4 * (int64_t)B.level.size()) + align_up(4 * (int64_t)B.desc.size());
    }
    const int64_t nX = (int64_t)(s.ch2d.size() + s.leaks.size() + s.syn.size());
    // the HHChannel2D kinds of the cell: objects with the same powers,
    // tables, axes and instant flags share one record
    std::vector<int> kindOf(s.ch2d.size(), 0);
    std::vector<size_t> kindRep;
    {
        std::map<std::vector<double>, int> kinds;
        for (size_t k = 0; k < s.ch2d.size(); ++k) {
            const HostChan2D& h = s.ch2d[k];
            std::vector<double> key(h.pw, h.pw + 3);
            for (int a = 0; a < 3; ++a) { key.push_back(h.tabA[a]); key.push_back(h.tabB[a]); }
            for (int q = 0; q < 6; ++q) key.push_back(h.dep[q]);
            key.push_back(h.instant);
            auto it = kinds.find(key);
            if (it == kinds.end()) {
                it = kinds.emplace(key, (int)kindRep.size()).first;
                kindRep.push_back(k);
            }
            kindOf[k] = it->second;
        }
    }
    std::vector<Item> items = {
        {&d.oV, 8 * n}, {&d.oInitVm, 8 * n}, {&d.oCmByDt, 8 * n}, {&d.oEmByRm, 8 * n}, {&d.oDbase, 8 * n}, {&d.oOff, 8 * n},
        {&d.oDefer, 8 * n},
        {&d.oChType, 4 * CS}, {&d.oGbar, gt * CS}, {&d.oEk, gt * CS}, {&d.oMod, 8 * CS}, {&d.oExtCa, 8 * CS},
        {&d.oState, 8 * GS},
        {&d.oPoolC, 8 * PS}, {&d.oPoolCa, 8 * PS}, {&d.oPoolBasal, 8 * PS}, {&d.oPoolF1, 8 * PS},
        {&d.oPoolF2, 8 * PS}, {&d.oPoolCeil, 8 * PS}, {&d.oPoolFloor, 8 * PS}, {&d.oPoolICa, 8 * PS},
        {&d.oPoolCount, 4 * n},
        {&d.oLevelStart, 4 * (int64_t)s.levelStart.size()}, {&d.oUnits, 4 * (int64_t)s.units.size()},
        {&d.oGrpMem, 4 * nMem}, {&d.oGrpInfo, 16 * nGroups}, {&d.oHJ0, 8 * (int64_t)s.HJ0.size()},
        // at least one descriptor: the step kernels load a level's
        // descriptors ahead without a branch (clamped to the last unit)
        {&d.oDesc, 4 * (int64_t)std::max<size_t>(s.udesc.size(), kDescInts)},
        {&d.oWork, workB},
        {&d.oSpkCompt, 4 * (int64_t)s.spkCompt.size()}, {&d.oSpkEdge, 4 * (int64_t)s.spkCompt.size()},
        {&d.oSpkThr, 8 * (int64_t)s.spkCompt.size()}, {&d.oSpkRefr, 8 * (int64_t)s.spkCompt.size()},
        {&d.oSpkLast, 8 * (int64_t)s.spkCompt.size()}, {&d.oSpkFired, 4 * (int64_t)s.spkCompt.size()},
        {&d.oSteps, 8},
        // message-passing solve (split mode)
        {&d.oMsgDesc, msgOn ? 4 * (int64_t)s.mdesc.size() : 0},
        {&d.oRowPos, msgOn ? 4 * (int64_t)n : 0},
        // rows in unit order, the root row, two dummy rows
        {&d.oDA, msgOn ? 8 * (2 * (int64_t)s.units.size() + 3) : 0},
        {&d.oBA, msgOn ? 8 * (2 * (int64_t)s.units.size() + 3) : 0},
        {&d.oFillU, msgOn ? 8 * ((int64_t)s.units.size() + 1) : 0},
        {&d.oVMid, msgOn ? 8 * ((int64_t)n + 1) : 0},
        // block form (large cells, split mode)
        {&d.oBlocks, blockBytes},
        // HHChannel2D objects
        {&d.o2dStart, s.ch2d.empty() ? 0 : 4 * ((int64_t)n + 1)},
        {&d.o2dChan, (int64_t)sizeof(Chan2DDev) * (int64_t)s.ch2d.size()},
        {&d.o2dType, (int64_t)sizeof(Chan2DType) * (int64_t)kindRep.size()},
        {&d.o2dState, 24 * (int64_t)s.ch2d.size()},
        {&d.o2dGk, 8 * (int64_t)s.ch2d.size()},
        {&d.o2dExt, nX == 0 ? 0 : 16 * (int64_t)n},
        // PulseGens
        {&d.oPulse, (int64_t)sizeof(PulseDev) * (int64_t)s.pulses.size()},
        {&d.oVClamp, (int64_t)sizeof(VClampDev) * (int64_t)s.vclamps.size()},
        {&d.oPulseStart, 4 * (int64_t)s.pulseStart.size()},
        {&d.oPulseIdx, 4 * (int64_t)s.pulseIdx.size()},
        {&d.oVcStart, 4 * (int64_t)s.vcStart.size()},
        {&d.oVcIdx, 4 * (int64_t)s.vcIdx.size()},
        {&d.oInjOf, 4 * (int64_t)n},
        // Leakage sends and SynChans
        {&d.oXFlag, nX == 0 ? 0 : 4 * (int64_t)n},
        {&d.oLeakStart, s.leaks.empty() ? 0 : 4 * ((int64_t)n + 1)},
        {&d.oLeak, (int64_t)sizeof(LeakDev) * (int64_t)s.leaks.size()},
        {&d.oSynStart, s.syn.empty() ? 0 : 4 * ((int64_t)n + 1)},
        {&d.oSyn, (int64_t)sizeof(SynDev) * (int64_t)s.syn.size()},
        {&d.oActCur, 4 * (int64_t)s.syn.size()},
        {&d.oActEnd, 4 * (int64_t)s.syn.size()},
    };
    int64_t total = 0;
    for (auto& it : items) total += align_up(std::max<int64_t>(it.bytes, 8));
    int64_t base = 0;
    int rc = arena_alloc(e, total, &base);
    if (rc) return rc;
    hc.region.off = base;
    hc.region.bytes = total;
    int64_t cur = base;
    for (auto& it : items) {
        *it.off = cur;
        cur += align_up(std::max<int64_t>(it.bytes, 8));
    }
    hc.bytes = total;

    d.n = n;
    d.probeStride = 1;
    d.nChanSlots = s.chanSlots;
    d.nGateSlots = s.gateSlots;
    d.nPoolSlots = s.poolSlots;
    d.hasModulation = s.hasModulation;
    d.hasPools = s.poolSlots > 0;
    d.caAdvance = s.caAdvance;
    d.nLevels = (int)s.levelStart.size() - 1;
    d.nUnits = (int)s.units.size();
    {
        // the trailing levels of one unit each, when there are at least two
        // (at most the last kMaxChain of them: their descriptors take LDS)
        int l = d.nLevels;
        while (l > 0 && s.levelStart[l] - s.levelStart[l - 1] == 1 && d.nLevels - l < kMaxChain) --l;
        d.chainLevel = d.nLevels - l >= 2 ? l : d.nLevels;
    }
    d.nHJ = (int)s.HJ0.size();
    d.nTerm = s.nTerm;
    d.nFill = s.nFill;
    d.nGroups = (int)nGroups;
    d.nSpike = (int)s.spkCompt.size();
    d.cellId = hc.id;
    d.live = 1;
    // engine type ids
    int tsid = 0;
    rc = table_set_id(e, s.tables, &tsid);
    if (rc) return rc;
    hc.tableSet = tsid;
    d.tableSet = tsid;
    hc.typeIds.clear();
    for (auto& t : s.localTypes) hc.typeIds.push_back(type_id(e, t, tsid));
    std::vector<int32_t> chType(CS, -1);
    for (int64_t k = 0; k < CS; ++k)
        chType[k] = s.chTypeLocal[k] >= 0 ? hc.typeIds[s.chTypeLocal[k]] : -1;
    d.uniform = s.uniform ? 1 : 0;
    for (int k = 0; k < kMaxSlots; ++k)
        d.slotType[k] = (k < s.chanSlots && s.uniform) ? chType[(int64_t)k * n] : -1;
    fill_fast_slots(e, d);

    // upload: the whole region in one copy (zero-filled, so zero arrays need
    // no staging)
    Stager st(base, total);
    auto put_gt = [&](int64_t off, const std::vector<double>& v) {
        if (e->fp32) st.put(off, std::vector<float>(v.begin(), v.end()));
        else st.put(off, v);
    };
    std::vector<int32_t> mem, info;
    for (auto& g : s.groups) {
        info.push_back((int32_t)mem.size());
        info.push_back((int32_t)g.mem.size());
        info.push_back(g.hjOff);
        info.push_back(g.deferLast);
        for (int x : g.mem) mem.push_back(x);
    }
    auto i32 = [](const std::vector<int>& v) { return std::vector<int32_t>(v.begin(), v.end()); };
    st.put(d.oV, s.V);
    st.put(d.oInitVm, s.initVm);
    st.put(d.oCmByDt, s.CmByDt);
    st.put(d.oEmByRm, s.EmByRm);
    st.put(d.oDbase, s.Dbase);
    st.put(d.oOff, s.Off);
    st.put(d.oDefer, i32(s.defer));
    st.put(d.oChType, chType);
    put_gt(d.oGbar, s.Gbar);
    put_gt(d.oEk, s.Ek);
    st.put(d.oMod, s.mod);
    st.put(d.oState, s.state);
    st.put(d.oPoolC, s.pC);
    st.put(d.oPoolCa, s.pCa);
    st.put(d.oPoolBasal, s.pBasal);
    st.put(d.oPoolF1, s.pF1);
    st.put(d.oPoolF2, s.pF2);
    st.put(d.oPoolCeil, s.pCeil);
    st.put(d.oPoolFloor, s.pFloor);
    st.put(d.oPoolCount, i32(s.poolCount));
    st.put(d.oLevelStart, i32(s.levelStart));
    st.put(d.oUnits, i32(s.units));
    st.put(d.oGrpMem, mem);
    st.put(d.oGrpInfo, info);
    st.put(d.oHJ0, s.HJ0);
    st.put(d.oDesc, s.udesc);
    st.put(d.oSpkCompt, i32(s.spkCompt));
    st.put(d.oSpkEdge, i32(s.spkEdge));
    st.put(d.oSpkThr, s.spkThr);
    st.put(d.oSpkRefr, s.spkRefr);
    if (msgOn) {
        st.put(d.oMsgDesc, s.mdesc);
        st.put(d.oRowPos, s.rowPos);
    }
    if (blockOn) {
        // the BlockDev records (the blocks, then the top), then each one's
        // arrays
        std::vector<BlockDev> recs(s.blocks.size());
        int64_t off = d.oBlocks + align_up((int64_t)sizeof(BlockDev) * (int64_t)s.blocks.size());
        int32_t maxLds = 0;
        for (size_t k = 0; k < s.blocks.size(); ++k) {
            const auto& B = s.blocks[k];
            BlockDev& r = recs[k];
            std::memset(&r, 0, sizeof r);
            r.nRows = B.nRows;
            r.nExt = B.nExt;
            r.nT = B.nT;
            r.nTExt = (int32_t)B.tExt.size();
            r.nTImp = (int32_t)B.tImp.size();
            r.nHJ = (int32_t)B.hj.size();
            r.nF = (int32_t)B.fills.size();
            r.nLevels = (int32_t)(B.level.size() - 1) / kBlockForms;
            r.nUnits = (int32_t)(B.desc.size() / kDescInts);
            r.ldsBytes = (int32_t)B.ldsBytes;
            if (k + 1 < s.blocks.size()) maxLds = std::max(maxLds, r.ldsBytes);
            r.oRows = off;  st.put(off, B.rows);  off += align_up(4 * (int64_t)B.rows.size() + 4);
            r.oTExt = off;  st.put(off, B.tExt);  off += align_up(4 * (int64_t)B.tExt.size() + 4);
            r.oTImp = off;  st.put(off, B.tImp);  off += align_up(4 * (int64_t)B.tImp.size() + 4);
            r.oFill = off;  st.put(off, B.fills); off += align_up(4 * (int64_t)B.fills.size() + 4);
            r.oHJ = off;    st.put(off, B.hj);    off += align_up(4 * (int64_t)B.hj.size() + 4);
            r.oLevel = off; st.put(off, B.level); off += align_up(4 * (int64_t)B.level.size());
            r.oDesc = off;  st.put(off, B.desc);  off += align_up(4 * (int64_t)B.desc.size());
        }
        st.put(d.oBlocks, recs);
        d.nBlocks = (int32_t)s.blocks.size() - 1;
        d.blockLds = maxLds;
        d.topLds = (int32_t)s.blocks.back().ldsBytes;
    }
    if (!s.ch2d.empty()) {
        std::vector<int> tid(s.tables2d.size(), -1);
        for (size_t k = 0; k < s.tables2d.size(); ++k)
            if ((rc = table2d_id(e, s.tables2d[k], &tid[k]))) return rc;
        std::vector<Chan2DType> ty(kindRep.size());
        for (size_t t = 0; t < kindRep.size(); ++t) {
            const HostChan2D& h = s.ch2d[kindRep[t]];
            Chan2DType& y = ty[t];
            std::memset(&y, 0, sizeof y);
            for (int a = 0; a < 3; ++a) {
                y.pw[a] = h.pw[a];
                y.tabA[a] = h.tabA[a] >= 0 ? tid[h.tabA[a]] : -1;
                y.tabB[a] = h.tabB[a] >= 0 ? tid[h.tabB[a]] : -1;
                y.pair[a] = -1;
                if (y.tabA[a] >= 0 && y.tabB[a] >= 0 &&
                    (rc = pair2d_id(e, y.tabA[a], y.tabB[a], s.tables2d[h.tabA[a]], s.tables2d[h.tabB[a]],
                                    &y.pair[a])))
                    return rc;
            }
            for (int q = 0; q < 6; ++q) y.dep[q] = h.dep[q];
            y.instant = h.instant;
        }
        std::vector<Chan2DDev> ch(s.ch2d.size());
        std::vector<double> stt(3 * s.ch2d.size());
        for (size_t k = 0; k < s.ch2d.size(); ++k) {
            const HostChan2D& h = s.ch2d[k];
            Chan2DDev& c = ch[k];
            c.Gbar = h.Gbar; c.Ek = h.Ek; c.mod = h.mod;
            for (int a = 0; a < 3; ++a) stt[3 * k + a] = h.st[a];
            c.type = kindOf[k];
            c.conc1 = h.conc1; c.conc2 = h.conc2;
            c.compt = h.compt;
        }
        st.put(d.o2dStart, s.ch2dStart);
        st.put(d.o2dChan, ch);
        st.put(d.o2dType, ty);
        st.put(d.o2dState, stt);
    }
    if (!s.pulses.empty()) {
        std::vector<PulseDev> pd(s.pulses.size());
        for (size_t k = 0; k < s.pulses.size(); ++k) {
            const HostPulse& h = s.pulses[k];
            PulseDev& q = pd[k];
            std::memset(&q, 0, sizeof q);
            q.compt = h.compt;
            q.entry = h.entry;
            q.n = h.n;
            q.trigMode = h.trigMode;
            q.src = h.src;
            q.srcBefore = h.src >= 0 && h.src < (int)k;
            q.base = h.base;
            q.out = h.base;
            q.trigTime = -1.0;
            for (int j = 0; j < h.n; ++j) {
                q.delay[j] = h.delay[j];
                q.width[j] = h.width[j];
                q.level[j] = h.level[j];
            }
        }
        st.put(d.oPulse, pd);
    }
    st.put(d.oPulseStart, s.pulseStart);
    st.put(d.oPulseIdx, s.pulseIdx);
    st.put(d.oVcStart, s.vcStart);
    st.put(d.oVcIdx, s.vcIdx);
    if (!s.vclamps.empty()) {
        std::vector<VClampDev> vd(s.vclamps.size());
        for (size_t k = 0; k < s.vclamps.size(); ++k) {
            VClampDev& q = vd[k];
            std::memset(&q, 0, sizeof q);
            q.compt = s.vclamps[k].compt;
            q.entry = s.vclamps[k].entry;
            q.mode = s.vclamps[k].mode;
            q.Kp = s.vclamps[k].gain;
        }
        st.put(d.oVClamp, vd);
    }
    if (nX > 0) {
        std::vector<int32_t> flag(n, 0);
        for (int i = 0; i < n; ++i)
            flag[i] = (!s.ch2d.empty() && s.ch2dStart[i] < s.ch2dStart[i + 1]) ||
                      (!s.leaks.empty() && s.leakStart[i] < s.leakStart[i + 1]) ||
                      (!s.syn.empty() && s.synStart[i] < s.synStart[i + 1]);
        st.put(d.oXFlag, flag);
    }
    if (!s.leaks.empty()) {
        st.put(d.oLeakStart, s.leakStart);
        st.put(d.oLeak, leak_records(s));
    }
    if (!s.syn.empty()) {
        st.put(d.oSynStart, s.synStart);
        st.put(d.oSyn, syn_records(s, hc.synConst, hc.synState));
    }
    if ((rc = st.flush(e))) return rc;
    d.nXObj = (int32_t)nX;
    d.nLeak = (int32_t)s.leaks.size();
    d.nSyn = (int32_t)s.syn.size();
    d.oAct = 0;
    d.msg = msgOn ? 1 : 0;
    d.n2d = (int32_t)s.ch2d.size();
    // one HHChannel2D on every compartment, in compartment order, of one
    // kind, and nothing else beside the solver: the external kernel's
    // direct path
    d.uni2d = 0;
    if (d.n2d == n && kindRep.size() == 1 && s.leaks.empty() && s.syn.empty()) {
        bool inOrder = true;
        for (int i = 0; i <= n && inOrder; ++i) inOrder = s.ch2dStart[i] == i;
        d.uni2d = inOrder ? 1 : 0;
    }
    d.nPulse = (int32_t)s.pulses.size();
    d.nVClamp = (int32_t)s.vclamps.size();
    d.nSetupEntries = s.nSetupEntries;
    d.values2dSent = 0;
    d.nSlots = s.nSlots;
    d.rootSlot = s.rootSlot;
    d.rootCnt = s.rootCnt;
    d.slowHJ = s.anySlowGroup ? 1 : 0;
    // work arrays (and the level table) in LDS when they fit, and the
    // groups' junction coefficients with them when those fit too
    const int64_t wl = work_lds_bytes(d), lb = level_bytes(d);
    d.hjInLds = 0;
    if (msgOn) {
        d.workInLds = 1;
        d.ldsBytes = (int32_t)msgLds;
    } else if (wl + lb <= e->ldsLimit) {
        d.workInLds = 1;
        d.hjInLds = s.anySlowGroup && wl + 8 * (int64_t)d.nHJ + lb <= e->ldsLimit ? 1 : 0;
        d.ldsBytes = (int32_t)align_up(wl + hj_lds_bytes(d) + lb);
    } else {
        d.workInLds = 0;   // work arrays and level table read from HBM
        d.ldsBytes = 0;
    }
    return 0;
}

int write_injections(hsg_engine* e, HostCell& hc) {
    const int m = (int)hc.injIdx.size();
    int64_t need = std::max(1, m) * (4 + 8 + 8) + 3 * kAlign;
    if (hc.inj.off < 0 || hc.inj.bytes < need) {
        if (hc.inj.off >= 0) e->arena.release(hc.inj.off, hc.inj.bytes);
        int64_t off;
        int rc = arena_alloc(e, need * 2, &off);
        if (rc) return rc;
        hc.inj.off = off;
        hc.inj.bytes = need * 2;
    }
    hc.dev.oInjIdx = hc.inj.off;
    hc.dev.oInjBasal = align_up(hc.inj.off + 4 * (int64_t)std::max(1, m));
    hc.dev.oInjVar = align_up(hc.dev.oInjBasal + 8 * (int64_t)std::max(1, m));
    std::vector<int32_t> idx(hc.injIdx.begin(), hc.injIdx.end());
    // compartment -> entry, for the split kernels' per-compartment lookup
    std::vector<int32_t> of(hc.s.n, -1);
    for (int k = 0; k < m; ++k) of[hc.injIdx[k]] = k;
    int rc;
    if ((rc = upload(e, hc.dev.oInjIdx, idx)) || (rc = upload(e, hc.dev.oInjBasal, hc.injBasal)) ||
        (rc = upload(e, hc.dev.oInjVar, hc.injVar)) || (rc = upload(e, hc.dev.oInjOf, of)))
        return rc;
    hc.dev.nInj = m;
    hc.injDirty = false;
    e->tablesDirty = true;
    return 0;
}

int write_ext(hsg_engine* e, HostCell& hc) {
    const int m = (int)hc.ext.size();
    if (m == 0) {
        if (hc.dev.nExt != 0) e->tablesDirty = true;
        hc.dev.nExt = 0;
        return 0;
    }
    int64_t need = (int64_t)m * 24 + 3 * kAlign;
    if (hc.extBlk.off < 0 || hc.extBlk.bytes < need) {
        if (hc.extBlk.off >= 0) e->arena.release(hc.extBlk.off, hc.extBlk.bytes);
        int64_t off;
        int rc = arena_alloc(e, need * 2, &off);
        if (rc) return rc;
        hc.extBlk.off = off;
        hc.extBlk.bytes = need * 2;
    }
    hc.dev.oExtIdx = hc.extBlk.off;
    hc.dev.oExtG = align_up(hc.extBlk.off + 4 * (int64_t)m);
    hc.dev.oExtGE = align_up(hc.dev.oExtG + 8 * (int64_t)m);
    std::vector<int32_t> idx;
    std::vector<double> g, ge;
    for (auto& kv : hc.ext) {
        idx.push_back(kv.first);
        g.push_back(kv.second.first);
        ge.push_back(kv.second.second);
    }
    int rc;
    if ((rc = upload(e, hc.dev.oExtIdx, idx)) || (rc = upload(e, hc.dev.oExtG, g)) ||
        (rc = upload(e, hc.dev.oExtGE, ge)))
        return rc;
    hc.dev.nExt = m;
    e->tablesDirty = true;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// helpers shared by the C ABI
// ---------------------------------------------------------------------------
namespace {

template <typename T>
int download(hsg_engine* e, int64_t off, size_t count, T* out) {
    if (!count) return 0;
    HIPCHK(hipMemcpyAsync(out, e->arena.base + off, count * sizeof(T), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

template <typename T>
int put1(hsg_engine* e, int64_t off, size_t index, T v) {
    HIPCHK(hipMemcpyAsync(e->arena.base + off + index * sizeof(T), &v, sizeof(T), hipMemcpyHostToDevice,
                          e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

HostCell* get_cell(hsg_engine* e, int32_t cell) {
    if (!e || cell < 0 || cell >= (int32_t)e->cells.size() || !e->cells[cell] || !e->cells[cell]->live)
        return nullptr;
    return e->cells[cell].get();
}

int put_gt(hsg_engine* e, int64_t off, size_t index, double v) {
    if (e->fp32) return put1<float>(e, off, index, (float)v);
    return put1<double>(e, off, index, v);
}

template <typename GT>
GT host_power(GT x, double p) {
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p == 3.0) return x * x * x;
    if (p == 4.0) { GT y = x * x; return y * y; }
    if (x > (GT)0) return (GT)std::exp((double)p * std::log((double)x));
    return (GT)0;
}

// Gk of one channel from the current gate states, computed the way the kernel
// does (ChannelStruct::process, hsolve/HSolveStruct.cpp:53-65).
template <typename GT>
double chan_gk(const HostCell& hc, int dk, const std::vector<double>& state) {
    const CellSetup& s = hc.s;
    const int n = s.n, i = s.chanCompt[dk], k = s.chanSlot[dk];
    const size_t o = (size_t)k * n + i;
    GT fraction = hc.dev.hasModulation ? (GT)s.mod[o] : (GT)1.0;
    int js = s.chanState[dk];
    if (s.chanXp[dk] > 0.0) fraction *= host_power<GT>((GT)state[(size_t)(js++) * n + i], s.chanXp[dk]);
    if (s.chanYp[dk] > 0.0) fraction *= host_power<GT>((GT)state[(size_t)(js++) * n + i], s.chanYp[dk]);
    if (s.chanZp[dk] > 0.0) fraction *= host_power<GT>((GT)state[(size_t)(js++) * n + i], s.chanZp[dk]);
    return (double)((GT)s.Gbar[o] * fraction);
}

int make_bins(hsg_engine* e, const std::vector<int32_t>& cells, hsg_engine::Plan& plan);
void free_plan(hsg_engine::Plan& plan);

// Drops the plans of cell subsets (their cells' bins may have changed).
void drop_sub_plans(hsg_engine* e) {
    for (auto& sp : e->subPlans) free_plan(sp.second);
    e->subPlans.clear();
}

// Brings the device copies of the cell records and engine tables up to date
// before a launch: a copy of each list into its existing buffer when only
// contents changed; the launch plan is rebuilt only when the live set or a
// cell's bin changed (planDirty).  A captured graph holds the buffers'
// addresses, so it stays valid unless a buffer moves or the plan changes.
int upload_tables(hsg_engine* e) {
    // a call that allocates nothing uploads the records but leaves the
    // launch plan to the next allocating call
    const bool doPlan = e->planDirty && !e->noAlloc;
    if (!e->tablesDirty && !doPlan) return 0;
    const int nc = (int)e->cells.size();
    std::vector<CellDev> devs(std::max(1, nc));
    std::memset(devs.data(), 0, sizeof(CellDev) * devs.size());
    std::vector<int32_t> live;
    for (int i = 0; i < nc; ++i) {
        if (!e->cells[i] || !e->cells[i]->live) continue;
        devs[i] = e->cells[i]->dev;
        live.push_back(i);
    }
    if (doPlan) {
        std::stable_sort(live.begin(), live.end(), [&](int a, int b) { return devs[a].n > devs[b].n; });
        e->order = live;
    }
    if (e->noAlloc && ((int)devs.size() > e->dCellsCap || std::max<int>(1, (int)e->types.size()) > e->dTypesCap ||
                       std::max<int>(1, (int)e->tsets.size()) > e->dTsetsCap ||
                       std::max<int>(1, (int)e->tabs2d.size()) > e->dTabs2dCap))
        return fail(HSG_E_STATE, "the engine's cell and type tables must grow, and this call allocates no device "
                                 "memory: step once with hsg_step or hsg_step_cells after adding cells");
    const void* before[4] = {e->dCells, e->dTypes, e->dTsets, e->dTabs2d};
    int rc;
    if ((rc = grow_dev(&e->dCells, &e->dCellsCap, (int)devs.size())) ||
        (rc = grow_dev(&e->dTypes, &e->dTypesCap, std::max<int>(1, (int)e->types.size()))) ||
        (rc = grow_dev(&e->dTsets, &e->dTsetsCap, std::max<int>(1, (int)e->tsets.size()))) ||
        (rc = grow_dev(&e->dTabs2d, &e->dTabs2dCap, std::max<int>(1, (int)e->tabs2d.size()))))
        return rc;
    const void* after[4] = {e->dCells, e->dTypes, e->dTsets, e->dTabs2d};
    if (std::memcmp(before, after, sizeof before) != 0) e->graphValid = false;
    HIPCHK(hipMemcpyAsync(e->dCells, devs.data(), sizeof(CellDev) * devs.size(), hipMemcpyHostToDevice, e->stream));
    if (!e->types.empty())
        HIPCHK(hipMemcpyAsync(e->dTypes, e->types.data(), sizeof(ChanType) * e->types.size(),
                              hipMemcpyHostToDevice, e->stream));
    if (!e->tsets.empty())
        HIPCHK(hipMemcpyAsync(e->dTsets, e->tsets.data(), sizeof(TableSetDev) * e->tsets.size(),
                              hipMemcpyHostToDevice, e->stream));
    if (!e->tabs2d.empty())
        HIPCHK(hipMemcpyAsync(e->dTabs2d, e->tabs2d.data(), sizeof(Table2DDev) * e->tabs2d.size(),
                              hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (doPlan) {
        drop_sub_plans(e);
        if ((rc = make_bins(e, e->order, e->plan))) return rc;
        e->planDirty = false;
        e->graphValid = false;
    }
    e->tablesDirty = false;
    return 0;
}

// Packs the slot types of a uniform cell for the fast channel pass:
// gate-state slots numbered in slot order, X, Y, Z.
void fill_fast_slots(hsg_engine* e, CellDev& d) {
    std::memset(d.fast, 0, sizeof d.fast);
    d.fastOk = d.uniform && d.nChanSlots > 0;
    int js = 0;
    for (int k = 0; k < kMaxSlots && d.fastOk; ++k) {
        const int t = d.slotType[k];
        if (t < 0) continue;
        const ChanType& ct = e->types[t];
        const double pw[3] = {ct.Xpower, ct.Ypower, ct.Zpower};
        const int col[3] = {ct.colX, ct.colY, ct.colZ};
        int code[3] = {0, 0, 0}, gidx[3] = {0, 0, 0}, cl[3] = {0, 0, 0};
        for (int ax = 0; ax < 3; ++ax) {
            if (!(pw[ax] > 0.0)) continue;
            if (pw[ax] != 1.0 && pw[ax] != 2.0 && pw[ax] != 3.0 && pw[ax] != 4.0) d.fastOk = 0;
            code[ax] = (int)pw[ax];
            cl[ax] = col[ax];
            gidx[ax] = js++;
        }
        if (js > 31 || cl[0] > 0xffff || cl[1] > 0xffff || cl[2] > 0xffff || ct.caDepLocal > 14) d.fastOk = 0;
        d.fast[k][0] = cl[0] | (cl[1] << 16);
        d.fast[k][1] = (int32_t)((uint32_t)cl[2] | (uint32_t)gidx[0] << 16 | (uint32_t)gidx[1] << 21 |
                                 (uint32_t)gidx[2] << 26 | 1u << 31);
        d.fast[k][2] = code[0] | code[1] << 3 | code[2] << 6 | (ct.instant & 7) << 9 | (ct.caDepLocal + 1) << 12;
    }
    if (!d.fastOk) std::memset(d.fast, 0, sizeof d.fast);
}

// LDS classes of the per-cell bins: a class's cells need at most
// kCuShare[k-1] and more than kCuShare[k] bytes (3, 4, 5, 6, 8 and 10 per CU)
constexpr int kLdsClasses = 5;
constexpr int64_t kCuShare[kLdsClasses] = {40960, 32768, 27306, 20480, 16384};

// Threads per cell workgroup: by cell size, and no more than the waves per
// CU of the fused kernels (12: three per SIMD) leave each of the
// workgroups that the bin's LDS lets reside on a CU.
int default_threads(int n, int lds) {
    const int byN = n >= 1024 ? 256 : n >= 256 ? 128 : 64;
    if (lds <= 0) return byN;
    const int resident = std::max(1, (160 * 1024) / lds);
    const int waves = std::max(1, 12 / resident);
    return std::min(byN, std::min(4, waves) * 64);
}

// Splits cells (largest first) into bins by kernel form and LDS need; each
// bin is one launch (per round or per tick) on its own stream, with its own
// dynamic LDS size and workgroup size.
void free_plan(hsg_engine::Plan& plan) {
    if (plan.dBuf && !plan.external) hipFree(plan.dBuf);
    plan.dBuf = nullptr;
    plan.cap = 0;
    plan.bins.clear();
}

// The bins of `cells` (host side only: cell lists and launch shapes; no
// device buffer is touched).
void plan_bins(hsg_engine* e, const std::vector<int32_t>& cells, std::vector<hsg_engine::Bin>& bins);
bool hh_pair(const CellDev& d);
bool ca_quartet(const CellDev& d);

// Bytes of the device tables of a plan's bins: per bin its cell order, then
// its channel tiles, each 256-byte aligned.
int64_t plan_bytes(hsg_engine* e, const std::vector<hsg_engine::Bin>& bins) {
    int64_t total = 0;
    for (const auto& b : bins) {
        int64_t nt = 0, nb = 0;
        for (int id : b.cells) {
            nt += (e->cells[id]->dev.n + kChannelTile - 1) / kChannelTile;
            nb += b.blocks ? e->cells[id]->dev.nBlocks : 0;
        }
        total += align_up(4 * (int64_t)std::max<size_t>(1, b.cells.size())) + align_up(8 * std::max<int64_t>(1, nt)) +
                 align_up(8 * std::max<int64_t>(1, nb));
    }
    return std::max<int64_t>(total, kAlign);
}

// Writes the bins' device tables into dBuf (plan_bytes(bins) bytes, device
// memory) with one copy, and points every bin at its part.
int write_plan_tables(hsg_engine* e, std::vector<hsg_engine::Bin>& bins, char* dBuf, int64_t bytes) {
    std::vector<char> img((size_t)bytes, 0);
    int64_t off = 0;
    for (auto& b : bins) {
        std::memcpy(img.data() + off, b.cells.data(), 4 * b.cells.size());
        b.dOrder = reinterpret_cast<int32_t*>(dBuf + off);
        off += align_up(4 * (int64_t)std::max<size_t>(1, b.cells.size()));
        // channel-kernel tiles: (cell, first compartment), kChannelTile each
        std::vector<int2> tiles;
        for (int id : b.cells)
            for (int i0 = 0; i0 < e->cells[id]->dev.n; i0 += kChannelTile) tiles.push_back(make_int2(id, i0));
        b.nTiles = (int)tiles.size();
        if (!tiles.empty()) std::memcpy(img.data() + off, tiles.data(), sizeof(int2) * tiles.size());
        b.dTiles = reinterpret_cast<int2*>(dBuf + off);
        off += align_up(8 * std::max<int64_t>(1, (int64_t)tiles.size()));
        // block-kernel tiles: (cell, block)
        std::vector<int2> btiles;
        if (b.blocks)
            for (int id : b.cells)
                for (int k = 0; k < e->cells[id]->dev.nBlocks; ++k) btiles.push_back(make_int2(id, k));
        b.nBlkTiles = (int)btiles.size();
        if (!btiles.empty()) std::memcpy(img.data() + off, btiles.data(), sizeof(int2) * btiles.size());
        b.dBlkTiles = reinterpret_cast<int2*>(dBuf + off);
        off += align_up(8 * std::max<int64_t>(1, (int64_t)btiles.size()));
    }
    if (off > bytes) return fail(HSG_E_STATE, "launch plan larger than its buffer");
    HIPCHK(hipMemcpyAsync(dBuf, img.data(), (size_t)off, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

// The plan of `cells`: bins, then their device tables in the plan's own
// buffer (grown by one allocation when too small).
int make_bins(hsg_engine* e, const std::vector<int32_t>& cells, hsg_engine::Plan& plan) {
    plan_bins(e, cells, plan.bins);
    if (plan.bins.empty()) return 0;
    const int64_t need = plan_bytes(e, plan.bins);
    if (plan.cap < need) {
        if (plan.dBuf && !plan.external) HIPCHK(hipFree(plan.dBuf));
        plan.dBuf = nullptr;
        plan.cap = 0;
        plan.external = false;
        HIPCHK(hipMalloc(&plan.dBuf, (size_t)need));
        plan.cap = need;
    }
    return write_plan_tables(e, plan.bins, plan.dBuf, plan.cap);
}

void plan_bins(hsg_engine* e, const std::vector<int32_t>& cells, std::vector<hsg_engine::Bin>& bins) {
    bins.clear();
    if (cells.empty()) return;
    // split mode: message-passing cells by workgroup-size class (largest
    // cell >= 1024: 256 threads, >= 256: 128, else 64); other cells (and
    // every cell in fused mode) by LDS need (> 32 KB, > 16 KB, the rest)
    // cells with HHChannel2D objects go to bins of their own (6, 7), which
    // run one tick per launch after that tick's 2D-channel kernel
    // cells with junction groups of four or more members (slowHJ) in bins of
    // their own too (8-15), launched with the kernels that carry the register
    // form of those groups
    // and, in automatic mode, large cells without the message-passing form
    // in split bins of their own (16-31)
    const bool autoMode = e->opt.kernel_mode == HSG_KERNEL_AUTO;
    // the fine LDS classes unless they would need more bins than the engine
    // has streams (pairs double a split bin): then one class per kind
    bool fine = true;
    auto cls = [&](int id) {
        const CellDev& d = e->cells[id]->dev;
        const int s = (d.slowHJ ? 8 : 0) + (autoMode && !d.msg && d.n >= kAutoSplitMin ? 16 : 0);
        if (d.n2d > 0) return s + (d.msg ? 6 : 7);
        if (d.msg) return s + (d.n >= 1024 ? 0 : d.n >= 256 ? 1 : 2);
        // the rest by how many of their workgroups' LDS fit a CU's 160 KiB:
        // three (more than 40 KiB), four, five, six, eight, ten or more (a
        // bin's dynamic LDS is its largest cell's: fine classes waste little)
        const int64_t wk = work_lds_bytes(d) + hj_lds_bytes(d) + level_bytes(d);
        int sub = 0;
        while (fine && sub < kLdsClasses && wk <= kCuShare[sub]) ++sub;
        return s + 3 + (sub << 8);
    };
    {
        std::set<int> kinds;
        for (int id : cells) kinds.insert(cls(id));
        if (2 * kinds.size() > (size_t)hsg_engine::kMaxBins) fine = false;
    }
    std::vector<int32_t> sorted(cells);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int a, int b) {
        const int ca = cls(a), cb = cls(b);
        if (ca != cb) return ca < cb;
        return e->cells[a]->dev.n > e->cells[b]->dev.n;
    });
    hsg_engine::Bin cur;
    int curCls = -1;
    for (int id : sorted) {
        const int k = cls(id);
        if (k != curCls && !cur.cells.empty()) {
            bins.push_back(cur);
            cur = hsg_engine::Bin();
        }
        curCls = k;
        cur.key = k;
        cur.msg = (k & 7) < 3 || (k & 7) == 6;
        cur.hasExt = (k & 7) >= 6;
        cur.slow = (k & 8) != 0;
        cur.split = e->split || cur.msg || (k & 16) != 0;
        cur.cells.push_back(id);
    }
    if (!cur.cells.empty()) bins.push_back(cur);
    // a split bin of few cells leaves most of the chip idle during its solve:
    // its two halves (cells alternating by size) run as a pipelined pair, the
    // bandwidth-bound kernels of one half beside the solve of the other
    for (size_t b = 0; b < bins.size() && bins.size() < (size_t)hsg_engine::kMaxBins; ++b) {
        const size_t nb = bins[b].cells.size();
        if (!bins[b].split || nb < 2 || nb > (size_t)kPairMaxCells) continue;
        hsg_engine::Bin second = bins[b];
        std::vector<int32_t> all;
        all.swap(bins[b].cells);
        second.cells.clear();
        for (size_t i = 0; i < nb; ++i) (i % 2 ? second.cells : bins[b].cells).push_back(all[i]);
        bins[b].pairNext = true;
        second.key = bins[b].key + 32;
        bins.insert(bins.begin() + b + 1, second);
        ++b;
    }
    for (auto& b : bins) {
        int maxN = 0, maxSlots = 0;
        bool allUniform = true;
        bool lean = true, hh = true, ca4 = true, uni2d = !b.cells.empty();
        for (int id : b.cells) {
            const CellDev& d = e->cells[id]->dev;
            b.lds = std::max<int>(b.lds, d.ldsBytes);
            maxN = std::max(maxN, d.n);
            maxSlots = std::max(maxSlots, d.nChanSlots);
            allUniform = allUniform && d.fastOk;
            lean = lean && d.nPulse == 0 && d.nVClamp == 0 && d.nXObj == 0 && !d.hasPools;
            hh = hh && hh_pair(d);
            ca4 = ca4 && ca_quartet(d);
            uni2d = uni2d && d.uni2d != 0;
        }
        b.uni2dAll = uni2d;
        b.sig = ca4 ? 2 : 0;
        b.lean = lean ? (hh ? 2 : 1) : 0;
        b.chanMS = allUniform ? (maxSlots <= 2 ? 2 : maxSlots <= 4 ? 4 : 8) : 0;
        b.sumN = 0;
        for (int id : b.cells) b.sumN += e->cells[id]->dev.n;
        // the per-cell kernels are built for at most 256 threads per
        // workgroup (wider ones measured no faster, 512 slower)
        b.threads = e->opt.block_threads > 0 ? std::max(64, std::min(256, (e->opt.block_threads + 63) / 64 * 64))
                                             : default_threads(maxN, b.split ? 0 : b.lds);
        // a shape the caller chose for this bin class (it measured it)
        const auto sh = e->shapes.find(b.key);
        if (sh != e->shapes.end()) {
            if (sh->second.threads > 0) b.threads = sh->second.threads;
            b.fastSel = sh->second.fast;
        }
        // large cells' levels hold hundreds of units: their split solve
        // (not the message-passing one) sweeps a level with 512 threads
        b.solveThreads = b.split && !b.msg && e->opt.block_threads <= 0 && (sh == e->shapes.end() || sh->second.threads <= 0) &&
                                 maxN >= kWideSolveMin ? 512 : b.threads;
        // the block form when every cell of a split bin has it
        b.blocks = b.split && !b.msg;
        b.blockLds = 0;
        b.topLds = 0;
        for (int id : b.cells) {
            const CellDev& d = e->cells[id]->dev;
            b.blocks = b.blocks && d.nBlocks > 0;
            b.blockLds = std::max(b.blockLds, (int)d.blockLds);
            b.topLds = std::max(b.topLds, (int)d.topLds);
        }
    }
}

LaunchArgs make_args(hsg_engine* e, const int32_t* order, int nsteps, int first) {
    LaunchArgs a{};
    a.arena = e->arena.base;
    a.cells = e->dCells;
    a.order = order;
    a.types = e->dTypes;
    a.tsets = e->dTsets;
    a.tabs2d = e->dTabs2d;
    a.fp32 = e->fp32 ? 1 : 0;
    a.spikes = e->dSpikes;
    a.spikeCount = e->dSpikeCount;
    a.spikeCap = kSpikeCap;
    a.nsteps = nsteps;
    a.firstOfCall = first;
    a.ldsBytes = 0;
    a.dt = e->dt;
    return a;
}

// flushes one-shot inputs of every live cell to the device
int flush_inputs(hsg_engine* e) {
    for (auto& up : e->cells) {
        if (!up || !up->live) continue;
        HostCell& hc = *up;
        int rc;
        if (hc.injDirty && (rc = write_injections(e, hc))) return rc;
        if ((rc = write_ext(e, hc))) return rc;
        if (!hc.ica.empty()) {
            for (auto& kv : hc.ica) {
                int p = kv.first;
                size_t o = (size_t)hc.s.poolSlot[p] * hc.s.n + hc.s.poolCompt[p];
                if ((rc = put1<double>(e, hc.dev.oPoolICa, o, kv.second))) return rc;
            }
        }
    }
    return 0;
}

// bookkeeping after a step call: one-shot inputs are consumed
void consume_inputs(hsg_engine* e, HostCell& hc, int64_t nsteps) {
    const int n = hc.s.n;
    if (hc.prevExtSet) {
        hc.prevExtG.assign(n, 0.0);
        hc.prevExtGE.assign(n, 0.0);
        hc.prevExtSet = false;
    }
    if (nsteps == 1)
        for (auto& kv : hc.ext) {
            hc.prevExtG[kv.first] = kv.second.first;
            hc.prevExtGE[kv.first] = kv.second.second;
            hc.prevExtSet = true;
        }
    if (!hc.ext.empty()) {
        hc.ext.clear();
        hc.dev.nExt = 0;
        e->tablesDirty = true;
    }
    hc.ica.clear();
    hc.gkOverride.clear();
    for (auto& v : hc.injVar) v = 0.0;
    hc.stepsDone += nsteps;
}

// SimpleSynHandler::vProcess (synapse/SimpleSynHandler.cpp:104-112) for the
// next nsteps ticks of each given cell: per SynChan, the ticks whose end time
// reaches queued events and the activation sum(weight / dt) each sends, in
// the order the queue pops them.  One list for all cells goes to the device
// in one copy (events, then per-SynChan first and end positions); the
// SynChan step reads it tick by tick.  Events arriving later (from spikes of
// these ticks) are due after them: hsg_step keeps each call within the
// shortest connection delay.
int write_activations(hsg_engine* e, const std::vector<HostCell*>& cells, const std::vector<int64_t>& nsteps) {
    const double dt = e->dt;
    std::vector<ActDev> acts;
    std::vector<int32_t> cur, end;
    std::vector<int64_t> base(cells.size());
    for (size_t c = 0; c < cells.size(); ++c) {
        HostCell& hc = *cells[c];
        base[c] = (int64_t)cur.size();
        const int64_t s0 = hc.stepsDone, s1 = hc.stepsDone + nsteps[c];
        for (size_t k = 0; k < hc.s.syn.size(); ++k) {
            cur.push_back((int32_t)acts.size());
            auto& q = hc.synQ[k];
            size_t p = 0;
            while (p < q.size()) {
                // the first tick of this call whose end time reaches event p
                const double tq = q[p].first;
                int64_t st = s0 + 1;
                const double est = std::ceil(tq / dt);
                if (est > (double)st) st = est > (double)(s1 + 1) ? s1 + 1 : (int64_t)est;
                while (st > s0 + 1 && tq <= dt * (double)(unsigned long long)(st - 1)) --st;
                while (st <= s1 && !(tq <= dt * (double)(unsigned long long)st)) ++st;
                if (st > s1) break;
                const double t = dt * (double)(unsigned long long)st;
                double a = 0.0;
                while (p < q.size() && q[p].first <= t) {
                    a += q[p].second / dt;
                    ++p;
                }
                if (a != 0.0) acts.push_back(ActDev{st, a});
            }
            q.erase(q.begin(), q.begin() + (std::ptrdiff_t)p);
            end.push_back((int32_t)acts.size());
        }
    }
    if (cur.empty()) return 0;
    // an empty list after an empty list: the device cursors already say so
    if (acts.empty() && e->actEmpty && e->actCells == cells.size()) return 0;
    const int64_t nA = (int64_t)acts.size(), nS = (int64_t)cur.size();
    const int64_t offCur = align_up(nA * (int64_t)sizeof(ActDev));
    const int64_t offEnd = offCur + align_up(4 * nS);
    const int64_t need = offEnd + align_up(4 * nS);
    if (e->actBlk.off < 0 || e->actBlk.bytes < need) {
        if (e->actBlk.off >= 0) e->arena.release(e->actBlk.off, e->actBlk.bytes);
        e->actBlk.off = -1;
        int64_t off;
        int rc = arena_alloc(e, 2 * need, &off);
        if (rc) return rc;
        e->actBlk.off = off;
        e->actBlk.bytes = 2 * need;
    }
    std::vector<char> blob((size_t)need, 0);
    if (nA) std::memcpy(blob.data(), acts.data(), sizeof(ActDev) * (size_t)nA);
    std::memcpy(blob.data() + offCur, cur.data(), 4 * (size_t)nS);
    std::memcpy(blob.data() + offEnd, end.data(), 4 * (size_t)nS);
    int rc = upload(e, e->actBlk.off, blob);
    if (rc) return rc;
    for (size_t c = 0; c < cells.size(); ++c) {
        CellDev& d = cells[c]->dev;
        const int64_t a = e->actBlk.off, ac = e->actBlk.off + offCur + 4 * base[c],
                      ae = e->actBlk.off + offEnd + 4 * base[c];
        if (d.oAct != a || d.oActCur != ac || d.oActEnd != ae) {
            d.oAct = a;
            d.oActCur = ac;
            d.oActEnd = ae;
            e->tablesDirty = true;
        }
    }
    e->actEmpty = acts.empty();
    e->actCells = cells.size();
    return 0;
}

// An event (time, weight) into a handler's queue: after the events of equal
// or earlier time (SimpleSynHandler::addSpike, :93-97).
void queue_event(std::vector<std::pair<double, double>>& q, double time, double weight) {
    auto it = std::upper_bound(q.begin(), q.end(), time,
                               [](double t, const std::pair<double, double>& x) { return t < x.first; });
    q.insert(it, {time, weight});
}

// Moves the spikes of the last launches from the device ring into the
// engine's buffer and delivers each along its connections
// (Synapse::addSpike, synapse/Synapse.cpp:104-113): the spike at t = dt *
// step becomes the event (t + delay, weight).  Spikes are delivered by
// step, then cell, then spike generator (the order the solvers send them).
int route_spikes(hsg_engine* e) {
    unsigned long long cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, e->dSpikeCount, sizeof cnt, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (cnt == 0) return 0;
    const unsigned long long have = std::min<unsigned long long>(cnt, kSpikeCap);
    std::vector<SpikeRec> recs(have);
    HIPCHK(hipMemcpyAsync(recs.data(), e->dSpikes, sizeof(SpikeRec) * have, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemsetAsync(e->dSpikeCount, 0, sizeof(unsigned long long), e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (cnt > have) return fail(HSG_E_STATE, "spike ring overflow between routed stretches");
    std::sort(recs.begin(), recs.end(), [](const SpikeRec& a, const SpikeRec& b) {
        if (a.step != b.step) return a.step < b.step;
        if (a.cell != b.cell) return a.cell < b.cell;
        return a.gen < b.gen;
    });
    for (const SpikeRec& r : recs) {
        e->spikeBuf.push_back(r);
        HostCell* src = get_cell(e, r.cell);
        if (!src) continue;
        auto it = e->conns.find({r.cell, src->s.spkDesc[r.gen]});
        if (it == e->conns.end()) continue;
        const double t = e->dt * (double)(unsigned long long)r.step;
        for (const auto& tg : it->second) {
            HostCell* dst = get_cell(e, tg.cell);
            if (!dst) continue;
            const int q = dst->s.synOfDesc[tg.synchan];
            const int j = dst->s.syn[q].firstSyn + tg.synapse;
            queue_event(dst->synQ[q], t + dst->s.synDelay[j], dst->s.synW[j]);
        }
    }
    return 0;
}

// Longest stretch of ticks whose spikes are all due after it: a spike at
// step s through delay d is due at the first step s' > s with
// dt * s + d <= dt * s', so stretches of floor(d_min / dt) ticks (at least
// one) never need a spike of their own.
int64_t routed_stretch(hsg_engine* e) {
    double dmin = -1.0;
    for (auto& kv : e->conns)
        for (const auto& tg : kv.second) {
            HostCell* dst = get_cell(e, tg.cell);
            if (!dst) continue;
            const int q = dst->s.synOfDesc[tg.synchan];
            const double d = dst->s.synDelay[dst->s.syn[q].firstSyn + tg.synapse];
            if (dmin < 0.0 || d < dmin) dmin = d;
        }
    if (dmin < 0.0) return INT64_MAX;
    const double k = std::floor(dmin / e->dt);
    return k < 1.0 ? 1 : k > 1e12 ? (int64_t)1e12 : (int64_t)k;
}

// The fused kernel's fast channel pass (two-slot uniform cells) holds more
// registers, so fewer workgroups share a CU: it pays off where each thread
// has many compartments to sweep (measured crossover on MI355X).
#ifndef HSG_FAST_FUSED_MIN
#define HSG_FAST_FUSED_MIN 6
#endif
// The squid HH pair on the fast pass: two live slots, m^3 h then n^4 (no Z
// gate, no instant gate), without modulation (the kernels' kSigHH).
bool hh_pair(const CellDev& d) {
    return d.fastOk && d.nChanSlots == 2 && !d.hasModulation && d.fast[0][1] < 0 && d.fast[1][1] < 0 &&
           d.fast[0][2] == (3 | 1 << 3) && d.fast[1][2] == 4;
}

// Na, Kdr, KA and CaL on the fast pass: four live slots m^3 h, n^4, a^3 b,
// s^2 (no Z gate, no instant gate), without modulation (kSigCa4).
bool ca_quartet(const CellDev& d) {
    const int w[4] = {3 | 1 << 3, 4, 3 | 1 << 3, 2};
    bool ok = d.fastOk && d.nChanSlots == 4 && !d.hasModulation;
    for (int k = 0; k < 4 && ok; ++k) ok = d.fast[k][1] < 0 && d.fast[k][2] == w[k];
    return ok;
}

int fused_fast(const hsg_engine::Bin& b) {
    if (b.cells.empty() || b.fastSel == 0) return 0;
    // Na/Kdr/KA/CaL cells (junction groups or not): hsolve_step_ca_kernel
    if (b.chanMS == 4 && b.sig == 2) return 4;
    if (b.chanMS != 2 || b.slow) return 0;
    if (b.fastSel == 1) return 2;
    const double perThread = (double)b.sumN / (double)b.cells.size() / (double)b.threads;
    return perThread >= HSG_FAST_FUSED_MIN ? 2 : 0;
}

// Bandwidth phase of one split tick: the HHChannel2D objects (bins that have
// them), then the channel kernel over the bin's tiles.  On the fast pass,
// bins whose cells all have one HHChannel2D per compartment (and nothing
// else beside the solver) advance it inside the channel kernel instead.
int split_wide(hsg_engine* e, const hsg_engine::Bin& bin, LaunchArgs& a, hipStream_t st) {
    const bool fast = bin.chanMS > 0 && bin.fastSel != 0;
    const bool inRow = fast && bin.hasExt && bin.uni2dAll;
    if (bin.hasExt && !inRow) HIPCHK(hsg_launch_external(&a, bin.dTiles, bin.nTiles, st));
    a.ext2dInRow = inRow ? 1 : 0;
    if (fast) HIPCHK(hsg_launch_channel_fast(&a, bin.dTiles, bin.nTiles, e->fp32, bin.chanMS, bin.sig, st));
    else HIPCHK(hsg_launch_channel(&a, bin.dTiles, bin.nTiles, e->fp32, st));
    a.ext2dInRow = 0;
    return 0;
}

// Solve phase of one split tick: the solve stops at VMid; V, Ca and the
// tick's outputs follow over the batch.
int split_solve(hsg_engine* e, const hsg_engine::Bin& bin, LaunchArgs& a, hipStream_t st) {
    const int nb = (int)bin.cells.size();
    a.deferD = 1;
    if (bin.msg) {
        HIPCHK(hsg_launch_msg(&a, nb, bin.threads, e->fp32, st));
    } else if (bin.blocks) {
        // blocks forward, the cells' tops, blocks backward
        if (bin.nBlkTiles) HIPCHK(hsg_launch_block(&a, bin.dBlkTiles, bin.nBlkTiles, bin.slow, 1, bin.blockLds, st));
        HIPCHK(hsg_launch_top(&a, nb, 256, bin.slow, bin.topLds, st));
        if (bin.nBlkTiles) HIPCHK(hsg_launch_block(&a, bin.dBlkTiles, bin.nBlkTiles, bin.slow, 0, bin.blockLds, st));
    } else {
        HIPCHK(hsg_launch_solve(&a, nb, bin.solveThreads, e->fp32, bin.slow, st));
    }
    HIPCHK(hsg_launch_update(&a, bin.dTiles, bin.nTiles, e->fp32, bin.sig, st));
    HIPCHK(hsg_launch_outputs(&a, nb, st));
    return 0;
}

// One round: every bin launched for `chunk` steps on its own stream, forked
// from and joined back into the engine stream. The halves A, B of a
// pipelined pair alternate their bandwidth phases through two events: B's
// phase of tick k follows A's, and A's phase of tick k+1 follows B's of tick
// k, so each half's solve overlaps the other half's channel work. Every
// cell still sees its own kernels in tick order.
int launch_round(hsg_engine* e, std::vector<hsg_engine::Bin>& bins, int chunk, int first, bool timed = false) {
    // one stream and one set of events per bin: the plan never holds more
    // bins than the engine has streams
    if (bins.size() > (size_t)hsg_engine::kMaxBins) return fail(HSG_E_STATE, "launch plan has more bins than streams");
    // every stream and event indexed below exists (created with the engine;
    // a null handle would silently mean the null stream), and a pair's
    // second half is a bin of this plan
    if (!e->evFork) return fail(HSG_E_STATE, "engine streams not set up");
    for (size_t b = 0; b < bins.size(); ++b) {
        if (!e->binStream[b] || !e->evJoin[b] || !e->evPairA[b] || !e->evPairB[b] ||
            (timed && (!e->evBin0[b] || !e->evBin1[b])))
            return fail(HSG_E_STATE, "engine streams not set up");
        if (bins[b].pairNext && bins[b].split && b + 1 >= bins.size())
            return fail(HSG_E_STATE, "pipelined pair without its second half");
    }
    HIPCHK(hipEventRecord(e->evFork, e->stream));
    for (size_t b = 0; b < bins.size(); ++b) {
        hipStream_t st = e->binStream[b];
        HIPCHK(hipStreamWaitEvent(st, e->evFork, 0));
        if (timed) HIPCHK(hipEventRecord(e->evBin0[b], st));
        const int nb = (int)bins[b].cells.size();
        const bool split = bins[b].split;
        if (split && bins[b].pairNext && b + 1 < bins.size()) {
            hipStream_t st2 = e->binStream[b + 1];
            HIPCHK(hipStreamWaitEvent(st2, e->evFork, 0));
            if (timed) HIPCHK(hipEventRecord(e->evBin0[b + 1], st2));
            const hsg_engine::Bin& A = bins[b];
            const hsg_engine::Bin& B = bins[b + 1];
            for (int k = 0; k < chunk; ++k) {
                LaunchArgs a = make_args(e, A.dOrder, 1, first && k == 0);
                a.ldsBytes = A.lds;
                if (k > 0) HIPCHK(hipStreamWaitEvent(st, e->evPairB[b], 0));
                if (int rc = split_wide(e, A, a, st)) return rc;
                HIPCHK(hipEventRecord(e->evPairA[b], st));
                if (int rc = split_solve(e, A, a, st)) return rc;
                LaunchArgs c = make_args(e, B.dOrder, 1, first && k == 0);
                c.ldsBytes = B.lds;
                HIPCHK(hipStreamWaitEvent(st2, e->evPairA[b], 0));
                if (int rc = split_wide(e, B, c, st2)) return rc;
                HIPCHK(hipEventRecord(e->evPairB[b], st2));
                if (int rc = split_solve(e, B, c, st2)) return rc;
            }
            if (timed) {
                HIPCHK(hipEventRecord(e->evBin1[b], st));
                HIPCHK(hipEventRecord(e->evBin1[b + 1], st2));
            }
            HIPCHK(hipEventRecord(e->evJoin[b], st));
            HIPCHK(hipStreamWaitEvent(e->stream, e->evJoin[b], 0));
            HIPCHK(hipEventRecord(e->evJoin[b + 1], st2));
            HIPCHK(hipStreamWaitEvent(e->stream, e->evJoin[b + 1], 0));
            ++b;
            continue;
        }
        if (split || bins[b].hasExt) {
            // tick by tick: the bandwidth phase and the solve, or the
            // HHChannel2D objects and one fused tick
            for (int k = 0; k < chunk; ++k) {
                LaunchArgs a = make_args(e, bins[b].dOrder, 1, first && k == 0);
                a.ldsBytes = bins[b].lds;
                if (!split) {
                    HIPCHK(hsg_launch_external(&a, bins[b].dTiles, bins[b].nTiles, st));
                    HIPCHK(hsg_launch_step(&a, nb, bins[b].threads, e->fp32, bins[b].slow, fused_fast(bins[b]),
                                           false, st));
                    continue;
                }
                if (int rc = split_wide(e, bins[b], a, st)) return rc;
                if (int rc = split_solve(e, bins[b], a, st)) return rc;
            }
        } else {
            LaunchArgs a = make_args(e, bins[b].dOrder, chunk, first);
            a.ldsBytes = bins[b].lds;
            HIPCHK(hsg_launch_step(&a, nb, bins[b].threads, e->fp32, bins[b].slow, fused_fast(bins[b]),
                                   bins[b].lean, st));
        }
        if (timed) HIPCHK(hipEventRecord(e->evBin1[b], st));
        HIPCHK(hipEventRecord(e->evJoin[b], st));
        HIPCHK(hipStreamWaitEvent(e->stream, e->evJoin[b], 0));
    }
    return 0;
}

// The plan of a subset of the live cells (hsg_step_cells with unequal step
// counts, or the cells left once some have finished): built once per subset
// and kept until the main plan changes, so a serving loop that steps the
// same cells quantum after quantum plans them once.
int sub_plan(hsg_engine* e, const std::vector<int32_t>& cells, hsg_engine::Plan** out) {
    for (size_t k = 0; k < e->subPlans.size(); ++k)
        if (e->subPlans[k].first == cells) {
            if (k + 1 != e->subPlans.size()) {
                auto hit = std::move(e->subPlans[k]);
                e->subPlans.erase(e->subPlans.begin() + (std::ptrdiff_t)k);
                e->subPlans.push_back(std::move(hit));
            }
            *out = &e->subPlans.back().second;
            return 0;
        }
    hsg_engine::Plan p;
    if (e->subPlans.size() >= hsg_engine::kMaxSubPlans) {
        // reuse the least recently used plan's buffer
        p = std::move(e->subPlans.front().second);
        e->subPlans.erase(e->subPlans.begin());
    }
    int rc = make_bins(e, cells, p);
    if (rc) {
        free_plan(p);
        return rc;
    }
    e->subPlans.emplace_back(cells, std::move(p));
    *out = &e->subPlans.back().second;
    return 0;
}

int run_launches(hsg_engine* e, const std::vector<int32_t>& order, const std::vector<int64_t>& steps) {
    // order: cells to run; steps[i]: steps for order[i]
    const int K = e->opt.steps_per_launch > 0 ? e->opt.steps_per_launch : 50;
    std::vector<int64_t> left(steps);
    bool whole = order.size() == e->order.size();
    if (whole) {
        std::vector<int32_t> a(order), b(e->order);
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        whole = a == b;
        for (size_t i = 0; i < steps.size() && whole; ++i) whole = steps[i] == steps[0];
    }
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
    auto event_pair = [&](size_t k) -> std::pair<hipEvent_t, hipEvent_t> {
        while (e->evStart.size() <= k) {
            hipEvent_t a, b;
            hipEventCreate(&a);
            hipEventCreate(&b);
            e->evStart.push_back(a);
            e->evStop.push_back(b);
        }
        return {e->evStart[k], e->evStop[k]};
    };
    bool first = true;
    int rc = 0;
    int timedSteps = 0;   // ticks of the last round timed per bin (whole set, no graph)
    for (auto& b : e->plan.bins) { b.roundMs = 0; b.roundSteps = 0; }
    while (true) {
        std::vector<int32_t> active;
        int64_t chunk = K;
        for (size_t i = 0; i < order.size(); ++i)
            if (left[i] > 0) { active.push_back(order[i]); chunk = std::min(chunk, left[i]); }
        if (active.empty()) break;
        std::vector<hsg_engine::Bin>* bins = &e->plan.bins;
        if (!whole) {
            hsg_engine::Plan* p = nullptr;
            if ((rc = sub_plan(e, active, &p))) break;
            bins = &p->bins;
        }
        // steady state of the whole set: replay a captured graph when asked
        const int G = e->opt.graph_chunk;
        bool useGraph = whole && !first && G > 0 && chunk == K;
        if (useGraph) {
            int64_t minLeft = INT64_MAX;
            for (size_t i = 0; i < order.size(); ++i) minLeft = std::min(minLeft, left[i]);
            useGraph = minLeft >= (int64_t)G * K;
        }
        auto ev = event_pair(evs.size());
        hipEventRecord(ev.first, e->stream);
        if (useGraph) {
            if (!e->graphValid || e->graphSteps != K) {
                if (e->graphExec) { hipGraphExecDestroy(e->graphExec); e->graphExec = nullptr; }
                hipGraph_t g;
                hipError_t he = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
                if (he != hipSuccess) { rc = fail(HSG_E_DEVICE, hipGetErrorString(he)); break; }
                for (int r = 0; r < G && !rc; ++r) rc = launch_round(e, e->plan.bins, K, 0);
                he = hipStreamEndCapture(e->stream, &g);
                if (rc) break;
                if (he != hipSuccess) { rc = fail(HSG_E_DEVICE, hipGetErrorString(he)); break; }
                he = hipGraphInstantiate(&e->graphExec, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                if (he != hipSuccess) { rc = fail(HSG_E_DEVICE, hipGetErrorString(he)); break; }
                e->graphValid = true;
                e->graphSteps = K;
            }
            hipError_t he = hipGraphLaunch(e->graphExec, e->stream);
            if (he != hipSuccess) { rc = fail(HSG_E_DEVICE, hipGetErrorString(he)); break; }
            for (size_t i = 0; i < order.size(); ++i) left[i] -= (int64_t)G * K;
            e->lastLaunches += G;
        } else {
            if ((rc = launch_round(e, *bins, (int)chunk, first ? 1 : 0, whole))) break;
            if (whole) timedSteps = (int)chunk;
            for (size_t i = 0; i < order.size(); ++i) if (left[i] > 0) left[i] -= chunk;
            e->lastLaunches += 1;
        }
        hipEventRecord(ev.second, e->stream);
        evs.push_back(ev);
        first = false;
    }
    hipError_t he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess && !rc) rc = fail(HSG_E_DEVICE, std::string("step: ") + hipGetErrorString(he));
    for (auto& ev : evs) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) e->lastTotalMs += ms;
    }
    if (!rc && timedSteps > 0)
        for (size_t b = 0; b < e->plan.bins.size(); ++b) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, e->evBin0[b], e->evBin1[b]) == hipSuccess) {
                e->plan.bins[b].roundMs = ms;
                e->plan.bins[b].roundSteps = timedSteps;
            }
        }
    return rc;
}

// run_launches for hsg_step_cells_noalloc: the same rounds (no graph), with
// the launch tables of each active set written into the caller-independent
// static area `buf` (cap bytes) instead of a plan buffer the library
// allocates.  A new active set's tables are copied on the engine stream
// after the previous round has joined it, so no round in flight reads them.
int run_launches_static(hsg_engine* e, const std::vector<int32_t>& order, const std::vector<int64_t>& steps,
                        char* buf, int64_t cap) {
    const int K = e->opt.steps_per_launch > 0 ? e->opt.steps_per_launch : 50;
    std::vector<int64_t> left(steps);
    hsg_engine::Plan p;
    p.dBuf = buf;
    p.cap = cap;
    p.external = true;
    std::vector<int32_t> activePrev;
    bool first = true;
    int rc = 0;
    HIPCHK(hipEventRecord(e->evStart[0], e->stream));
    while (true) {
        std::vector<int32_t> active;
        int64_t chunk = K;
        for (size_t i = 0; i < order.size(); ++i)
            if (left[i] > 0) { active.push_back(order[i]); chunk = std::min(chunk, left[i]); }
        if (active.empty()) break;
        if (active != activePrev) {
            plan_bins(e, active, p.bins);
            const int64_t need = plan_bytes(e, p.bins);
            if (need > cap) {
                rc = fail(HSG_E_ARG, "launch tables of " + std::to_string(need) + " bytes exceed the " +
                                         std::to_string(cap) + "-byte static area of hsg_step_cells_noalloc");
                break;
            }
            if ((rc = write_plan_tables(e, p.bins, buf, cap))) break;
            activePrev = active;
        }
        if ((rc = launch_round(e, p.bins, (int)chunk, first ? 1 : 0))) break;
        for (size_t i = 0; i < order.size(); ++i) if (left[i] > 0) left[i] -= chunk;
        e->lastLaunches += 1;
        first = false;
    }
    HIPCHK(hipEventRecord(e->evStop[0], e->stream));
    hipError_t he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess && !rc) rc = fail(HSG_E_DEVICE, std::string("step: ") + hipGetErrorString(he));
    float ms = 0;
    if (!rc && hipEventElapsedTime(&ms, e->evStart[0], e->evStop[0]) == hipSuccess) e->lastTotalMs += ms;
    return rc;
}

int reinit_cells_host(hsg_engine* e, HostCell& hc) {
    // the device state (V, Ca pools, spike generators, gates, PulseGen
    // outputs) is reset by hsolve_reinit_kernel; here only the host mirrors
    CellSetup& s = hc.s;
    for (auto& c : s.pC) c = 0.0;
    // PulseGen::reinit sends baseLevel (device/PulseGen.cpp:426-433); the
    // solver's reinit keeps injectVarying, so the first step adds it
    for (const HostPulse& p : s.pulses) {
        for (size_t k = 0; k < hc.injIdx.size(); ++k)
            if (hc.injIdx[k] == p.compt) {
                hc.injVar[k] += p.base;
                hc.injDirty = true;
            }
    }
    // VClamp::reinit (biophysics/VClamp.cpp:313-347, clock tick 0): the
    // constants at the clock's dt, the command filter and error history
    // restarted at the compartment's initVm; the current carries over
    if (!s.vclamps.empty()) {
        std::vector<VClampDev> vd(s.vclamps.size());
        int rc = download(e, hc.dev.oVClamp, vd.size() * sizeof(VClampDev) / sizeof(double),
                          reinterpret_cast<double*>(vd.data()));
        if (rc) return rc;
        const double dt = e->dt;
        for (size_t k = 0; k < s.vclamps.size(); ++k) {
            HostVClamp& v = s.vclamps[k];
            VClampDev& q = vd[k];
            if (v.ti == 0.0) v.ti = dt;
            if (v.td < 0.0) v.td = 0.0;
            if (v.tau == 0.0) v.tau = 5.0 * dt;
            q.expt = (dt / v.tau > 1e-15) ? std::exp(-dt / v.tau) : 1.0 - dt / v.tau;
            q.tauByDt = v.tau / dt;
            q.dtByTi = dt / v.ti;
            q.tdByDt = v.td / dt;
            if (v.gain == 0.0) v.gain = s.Cm[v.compt] / dt;
            q.Kp = v.gain;
            q.v1 = q.e1 = q.e2 = 0.0;
            q.command = q.cmdIn = q.oldCmdIn = s.initVm[v.compt];
        }
        if ((rc = upload(e, hc.dev.oVClamp, vd))) return rc;
    }
    // SynChan::vReinit at the clock's dt, handlers' queues dropped
    // (SimpleSynHandler::vReinit)
    if (!s.syn.empty()) {
        for (size_t k = 0; k < s.syn.size(); ++k) {
            hc.synConst[k] = syn_constants(s.syn[k], e->dt);
            hc.synState[k] = {0.0, 0.0, 0.0};
            hc.synQ[k].clear();
        }
        int rc = upload(e, hc.dev.oSyn, syn_records(s, hc.synConst, hc.synState));
        if (rc) return rc;
        std::vector<int32_t> zero(s.syn.size(), 0);
        if ((rc = upload(e, hc.dev.oActCur, zero)) || (rc = upload(e, hc.dev.oActEnd, zero))) return rc;
    }
    if (hc.probe.off >= 0) {
        hc.probeBase = 0;
        int rc = put1<int64_t>(e, hc.dev.oProbeBase, 0, 0);
        if (rc) return rc;
    }
    e->tablesDirty = true;
    hc.ext.clear();
    hc.ica.clear();
    if (hc.dev.nExt) hc.dev.nExt = 0;
    hc.hsolveDt = e->dt;
    hc.stepsDone = 0;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* hsg_last_error(void) { return g_err.c_str(); }

int hsg_create(const hsg_options* opts, hsg_engine** out) {
    if (!out) return fail(HSG_E_ARG, "null output pointer");
    *out = nullptr;
    std::unique_ptr<hsg_engine> e(new hsg_engine());
    if (opts) e->opt = *opts;
    if (e->opt.precision != HSG_FP64 && e->opt.precision != HSG_MIXED)
        return fail(HSG_E_ARG, "precision must be HSG_FP64 or HSG_MIXED");
    e->fp32 = e->opt.precision == HSG_MIXED;
    if (e->opt.kernel_mode < HSG_KERNEL_AUTO || e->opt.kernel_mode > HSG_KERNEL_SPLIT)
        return fail(HSG_E_ARG, "kernel_mode must be HSG_KERNEL_AUTO, HSG_KERNEL_FUSED or HSG_KERNEL_SPLIT");
    e->split = e->opt.kernel_mode == HSG_KERNEL_SPLIT;
    e->device = e->opt.device;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (e->device < 0 || e->device >= ndev) return fail(HSG_E_ARG, "no such HIP device");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    int lds = 0;
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device));
    // dynamic LDS per workgroup: the device maximum less the kernel's static
    // LDS (the staged channel types), rounded down
    e->ldsLimit = 0;
    for (int cand : {lds - 4096, 65536, 32768}) {
        if (cand <= 0) continue;
        if (hsg_set_max_lds(cand) == hipSuccess) { e->ldsLimit = cand; break; }
        (void)hipGetLastError();
    }
    if (e->ldsLimit == 0) return fail(HSG_E_DEVICE, "cannot configure dynamic LDS for the step kernel");
    for (int b = 0; b < hsg_engine::kMaxBins; ++b) {
        HIPCHK(hipStreamCreateWithFlags(&e->binStream[b], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&e->evJoin[b], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&e->evPairA[b], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&e->evPairB[b], hipEventDisableTiming));
        HIPCHK(hipEventCreate(&e->evBin0[b]));
        HIPCHK(hipEventCreate(&e->evBin1[b]));
    }
    HIPCHK(hipEventCreateWithFlags(&e->evFork, hipEventDisableTiming));
    {
        // one timing pair ahead of any step call (the no-allocation entry
        // point creates nothing), and the static launch-table area of this
        // device, looked up once per device with the device current
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        e->evStart.push_back(a);
        e->evStop.push_back(b);
        void* sp = nullptr;
        long long sb = 0;
        HIPCHK(hsg_scratch_address(e->device, &sp, &sb));
        e->scratch = static_cast<char*>(sp);
        e->scratchBytes = sb;
    }
    int64_t cap = e->opt.mem_budget > 0 ? e->opt.mem_budget : (int64_t)64 << 20;
    HIPCHK(hipMalloc(&e->arena.base, cap));
    e->arena.cap = cap;
    e->arena.freeList[0] = cap;
    HIPCHK(hipMalloc(&e->dSpikes, sizeof(SpikeRec) * kSpikeCap));
    HIPCHK(hipMalloc(&e->dSpikeCount, sizeof(unsigned long long)));
    HIPCHK(hipMemsetAsync(e->dSpikeCount, 0, sizeof(unsigned long long), e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    *out = e.release();
    return 0;
}

int hsg_destroy(hsg_engine* e) {
    if (!e) return 0;
    hipSetDevice(e->device);
    if (e->stream) hipStreamSynchronize(e->stream);
    if (e->graphExec) hipGraphExecDestroy(e->graphExec);
    for (auto ev : e->evStart) hipEventDestroy(ev);
    for (auto ev : e->evStop) hipEventDestroy(ev);
    hipFree(e->arena.base);
    hipFree(e->dCells);
    hipFree(e->dTypes);
    hipFree(e->dTsets);
    hipFree(e->dTabs2d);
    hipFree(e->dSpikes);
    hipFree(e->dSpikeCount);
    free_plan(e->plan);
    drop_sub_plans(e);
    for (int b = 0; b < hsg_engine::kMaxBins; ++b) {
        if (e->binStream[b]) hipStreamDestroy(e->binStream[b]);
        if (e->evJoin[b]) hipEventDestroy(e->evJoin[b]);
        if (e->evPairA[b]) hipEventDestroy(e->evPairA[b]);
        if (e->evPairB[b]) hipEventDestroy(e->evPairB[b]);
        if (e->evBin0[b]) hipEventDestroy(e->evBin0[b]);
        if (e->evBin1[b]) hipEventDestroy(e->evBin1[b]);
    }
    if (e->evFork) hipEventDestroy(e->evFork);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return 0;
}

int hsg_add_cell(hsg_engine* e, const hsg_cell_desc* desc, double dt, int32_t caAdvance, int32_t* cell_id) {
    if (!e || !desc || !cell_id) return fail(HSG_E_ARG, "null argument");
    hipSetDevice(e->device);
    std::unique_ptr<HostCell> hc(new HostCell());
    std::string err;
    int rc = setup_cell(desc, dt, caAdvance, hc->s, err);
    if (rc) return fail(rc, err);
    // find a free id
    int id = -1;
    for (int i = 0; i < (int)e->cells.size(); ++i)
        if (!e->cells[i]) { id = i; break; }
    if (id < 0) { id = (int)e->cells.size(); e->cells.emplace_back(); }
    hc->id = id;
    hc->hsolveDt = dt;
    // SynChans as the objects hold them before any reinit: constants at
    // their default dt_, X = Y = Gk = 0, no queued events
    for (const HostSyn& q : hc->s.syn) hc->synConst.push_back(syn_constants(q, 25.0e-6));
    hc->synState.assign(hc->s.syn.size(), {0.0, 0.0, 0.0});
    hc->synQ.assign(hc->s.syn.size(), {});
    rc = place_cell(e, *hc);
    if (rc) {
        if (hc->region.off >= 0) e->arena.release(hc->region.off, hc->region.bytes);
        return rc;
    }
    for (int i = 0; i < hc->s.n; ++i)
        if (hc->s.hasInject[i]) {
            hc->injIdx.push_back(i);
            hc->injBasal.push_back(hc->s.injBasal[i]);
            hc->injVar.push_back(0.0);
        }
    rc = write_injections(e, *hc);
    if (rc) {
        e->arena.release(hc->region.off, hc->region.bytes);
        return rc;
    }
    hc->prevExtG.assign(hc->s.n, 0.0);
    hc->prevExtGE.assign(hc->s.n, 0.0);
    hc->live = true;
    e->cells[id] = std::move(hc);
    e->tablesDirty = true;
    e->planDirty = true;
    *cell_id = id;
    return 0;
}

int hsg_remove_cell(hsg_engine* e, int32_t cell) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    hipSetDevice(e->device);
    hipStreamSynchronize(e->stream);
    e->arena.release(hc->region.off, hc->region.bytes);
    if (hc->inj.off >= 0) e->arena.release(hc->inj.off, hc->inj.bytes);
    if (hc->probe.off >= 0) e->arena.release(hc->probe.off, hc->probe.bytes);
    if (hc->extBlk.off >= 0) e->arena.release(hc->extBlk.off, hc->extBlk.bytes);
    if (hc->act.off >= 0) e->arena.release(hc->act.off, hc->act.bytes);
    // its spike connections go with it
    for (auto it = e->conns.begin(); it != e->conns.end();) {
        auto& v = it->second;
        v.erase(std::remove_if(v.begin(), v.end(), [&](const hsg_engine::SpikeTarget& t) { return t.cell == cell; }),
                v.end());
        if (it->first.first == cell || v.empty()) it = e->conns.erase(it);
        else ++it;
    }
    e->cells[cell].reset();
    e->tablesDirty = true;
    e->planDirty = true;
    return 0;
}

static int reinit_some(hsg_engine* e, const std::vector<int32_t>& ids) {
    int rc;
    for (int id : ids)
        if ((rc = reinit_cells_host(e, *e->cells[id]))) return rc;
    if ((rc = upload_tables(e))) return rc;
    if (ids.empty()) return 0;
    int32_t* dSub = nullptr;
    HIPCHK(hipMalloc(&dSub, sizeof(int32_t) * ids.size()));
    HIPCHK(hipMemcpyAsync(dSub, ids.data(), sizeof(int32_t) * ids.size(), hipMemcpyHostToDevice, e->stream));
    LaunchArgs a = make_args(e, dSub, 0, 0);
    // the HHChannel2D objects reinit before the solver, from the values it
    // last sent
    hipError_t he = hsg_launch_chan2d_reinit(&a, (int)ids.size(), e->stream);
    if (he == hipSuccess) he = hsg_launch_reinit(&a, (int)ids.size(), e->fp32, e->stream);
    hipError_t hs = hipStreamSynchronize(e->stream);
    hipFree(dSub);
    // HSolveActive::reinit ends with sendValues: from now on the HHChannel2D
    // objects see the solver's Vm and [Ca]
    for (int id : ids) {
        CellDev& d = e->cells[id]->dev;
        if (d.n2d > 0 && !d.values2dSent) {
            d.values2dSent = 1;
            e->tablesDirty = true;
        }
    }
    if (he != hipSuccess) return fail(HSG_E_DEVICE, std::string("reinit launch: ") + hipGetErrorString(he));
    if (hs != hipSuccess) return fail(HSG_E_DEVICE, std::string("reinit: ") + hipGetErrorString(hs));
    return 0;
}

int hsg_reinit(hsg_engine* e, double dt) {
    if (!e) return fail(HSG_E_ARG, "null engine");
    if (!(dt > 0.0)) return fail(HSG_E_ARG, "reinit: dt must be positive");
    hipSetDevice(e->device);
    if (dt != e->dt) e->graphValid = false;   // the graph's launches carry dt
    e->dt = dt;
    e->reinitDone = true;
    std::vector<int32_t> ids;
    for (auto& c : e->cells) if (c && c->live) ids.push_back(c->id);
    HIPCHK(hipMemsetAsync(e->dSpikeCount, 0, sizeof(unsigned long long), e->stream));
    e->spikeBuf.clear();
    return reinit_some(e, ids);
}

int hsg_reinit_cell(hsg_engine* e, int32_t cell) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    if (!(e->dt > 0.0)) return fail(HSG_E_STATE, "call hsg_reinit once before reinitialising single cells");
    hipSetDevice(e->device);
    return reinit_some(e, {cell});
}

int hsg_step(hsg_engine* e, int64_t nsteps) {
    if (!e) return fail(HSG_E_ARG, "null engine");
    if (nsteps < 0) return fail(HSG_E_ARG, "negative step count");
    if (!(e->dt > 0.0)) return fail(HSG_E_STATE, "call hsg_reinit before stepping");
    if (nsteps == 0) return 0;
    hipSetDevice(e->device);
    int rc;
    bool anySyn = false;
    for (auto& c : e->cells) if (c && c->live && !c->s.syn.empty()) anySyn = true;
    // with SynChans: stretches short enough that no spike of a stretch is
    // due within it, the handlers' activations prepared before each and its
    // spikes routed after it
    const int64_t L = anySyn ? routed_stretch(e) : nsteps;
    std::vector<HostCell*> synCells;
    for (auto& c : e->cells) if (c && c->live && !c->s.syn.empty()) synCells.push_back(c.get());
    e->lastLaunches = 0;
    e->lastTotalMs = 0;
    for (int64_t done = 0; done < nsteps;) {
        const int64_t m = std::min(nsteps - done, L);
        if ((rc = flush_inputs(e))) return rc;
        if (anySyn && (rc = write_activations(e, synCells, std::vector<int64_t>(synCells.size(), m)))) return rc;
        if ((rc = upload_tables(e))) return rc;
        std::vector<int64_t> steps(e->order.size(), m);
        if (!e->order.empty() && (rc = run_launches(e, e->order, steps))) return rc;
        for (auto& c : e->cells) if (c && c->live) consume_inputs(e, *c, m);
        if (!e->conns.empty() && (rc = route_spikes(e))) return rc;
        done += m;
    }
    return 0;
}

int hsg_step_cells(hsg_engine* e, const int32_t* cells, const int64_t* nsteps, int32_t n) {
    if (!e || (n > 0 && (!cells || !nsteps))) return fail(HSG_E_ARG, "null argument");
    if (!(e->dt > 0.0)) return fail(HSG_E_STATE, "call hsg_reinit before stepping");
    hipSetDevice(e->device);
    std::vector<int32_t> ord;
    std::vector<int64_t> st;
    std::vector<char> seen(e->cells.size(), 0);
    for (int i = 0; i < n; ++i) {
        if (!get_cell(e, cells[i])) return fail(HSG_E_ARG, "no such cell");
        if (seen[cells[i]]) return fail(HSG_E_ARG, "a cell is listed twice");
        seen[cells[i]] = 1;
        if (nsteps[i] < 0) return fail(HSG_E_ARG, "negative step count");
        if (nsteps[i] > 0) { ord.push_back(cells[i]); st.push_back(nsteps[i]); }
    }
    if (!e->conns.empty()) return fail(HSG_E_UNSUPPORTED, "cells joined by spike connections step with hsg_step");
    int rc;
    if ((rc = flush_inputs(e))) return rc;
    std::vector<HostCell*> synCells;
    std::vector<int64_t> synSteps;
    for (size_t i = 0; i < ord.size(); ++i)
        if (!e->cells[ord[i]]->s.syn.empty()) {
            synCells.push_back(e->cells[ord[i]].get());
            synSteps.push_back(st[i]);
        }
    if ((rc = write_activations(e, synCells, synSteps))) return rc;
    e->lastLaunches = 0;
    e->lastTotalMs = 0;
    if ((rc = upload_tables(e))) return rc;
    if (!ord.empty() && (rc = run_launches(e, ord, st))) return rc;
    for (size_t i = 0; i < ord.size(); ++i) consume_inputs(e, *e->cells[ord[i]], st[i]);
    return 0;
}

int hsg_step_cells_noalloc(hsg_engine* e, const int32_t* cells, const int64_t* nsteps, int32_t n) {
    if (!e || (n > 0 && (!cells || !nsteps))) return fail(HSG_E_ARG, "null argument");
    if (!(e->dt > 0.0)) return fail(HSG_E_STATE, "call hsg_reinit before stepping");
    if (!e->scratch) return fail(HSG_E_STATE, "no static launch-table area on this device");
    hipSetDevice(e->device);
    std::vector<int32_t> ord;
    std::vector<int64_t> st;
    std::vector<char> seen(e->cells.size(), 0);
    for (int i = 0; i < n; ++i) {
        if (!get_cell(e, cells[i])) return fail(HSG_E_ARG, "no such cell");
        if (seen[cells[i]]) return fail(HSG_E_ARG, "a cell is listed twice");
        seen[cells[i]] = 1;
        if (nsteps[i] < 0) return fail(HSG_E_ARG, "negative step count");
        if (nsteps[i] > 0) { ord.push_back(cells[i]); st.push_back(nsteps[i]); }
    }
    if (!e->conns.empty()) return fail(HSG_E_UNSUPPORTED, "cells joined by spike connections step with hsg_step");
    // every allocation path below fails instead of allocating
    struct Guard {
        hsg_engine* e;
        ~Guard() { e->noAlloc = false; }
    } guard{e};
    e->noAlloc = true;
    int rc;
    if ((rc = flush_inputs(e))) return rc;
    std::vector<HostCell*> synCells;
    std::vector<int64_t> synSteps;
    for (size_t i = 0; i < ord.size(); ++i)
        if (!e->cells[ord[i]]->s.syn.empty()) {
            synCells.push_back(e->cells[ord[i]].get());
            synSteps.push_back(st[i]);
        }
    if ((rc = write_activations(e, synCells, synSteps))) return rc;
    e->lastLaunches = 0;
    e->lastTotalMs = 0;
    if ((rc = upload_tables(e))) return rc;
    if (!ord.empty() && (rc = run_launches_static(e, ord, st, e->scratch, e->scratchBytes))) return rc;
    for (size_t i = 0; i < ord.size(); ++i) consume_inputs(e, *e->cells[ord[i]], st[i]);
    return 0;
}

int hsg_get_field(hsg_engine* e, int32_t cell, int32_t field, const int32_t* idx, int32_t n, double* out) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    if (n < 0 || (n > 0 && (!idx || !out))) return fail(HSG_E_ARG, "bad index list");
    if (n == 0) return 0;
    hipSetDevice(e->device);
    CellSetup& s = hc->s;
    const int nc = s.n;
    std::vector<double> V, state, pool;
    int rc;
    auto needV = [&]() -> int { if (V.empty()) { V.resize(nc); return download(e, hc->dev.oV, nc, V.data()); } return 0; };
    auto needState = [&]() -> int {
        if (state.empty() && s.gateSlots > 0) {
            state.resize((size_t)s.gateSlots * nc);
            return download(e, hc->dev.oState, state.size(), state.data());
        }
        return 0;
    };
    for (int q = 0; q < n; ++q) {
        const int i = idx[q];
        if (field >= 1 && field <= 8) {
            if (i < 0 || i >= (int)s.descToHines.size() || s.descToHines[i] < 0)
                return fail(HSG_E_ARG, "unknown compartment index");
            const int h = s.descToHines[i];
            switch (field) {
            case HSG_F_VM: if ((rc = needV())) return rc; out[q] = V[h]; break;
            case HSG_F_IM: {
                // HSolve::getIm (hsolve/HSolveInterface.cpp:161-182)
                if ((rc = needV()) || (rc = needState())) return rc;
                double Im = s.EmByRm[h] - V[h] / s.Rm[h];
                for (int dk = 0; dk < (int)s.chanCompt.size(); ++dk) {
                    if (s.chanCompt[dk] != h) continue;
                    auto ov = hc->gkOverride.find(dk);
                    double gk = ov != hc->gkOverride.end() ? ov->second
                              : e->fp32 ? chan_gk<float>(*hc, dk, state) : chan_gk<double>(*hc, dk, state);
                    double ek = e->fp32 ? (double)(float)s.Ek[(size_t)s.chanSlot[dk] * nc + h]
                                        : s.Ek[(size_t)s.chanSlot[dk] * nc + h];
                    Im += (ek - V[h]) * gk;
                }
                double eg = hc->prevExtG[h], ege = hc->prevExtGE[h];
                if ((!s.ch2d.empty() && s.ch2dStart[h] < s.ch2dStart[h + 1]) ||
                    (!s.leaks.empty() && s.leakStart[h] < s.leakStart[h + 1]) ||
                    (!s.syn.empty() && s.synStart[h] < s.synStart[h + 1])) {
                    // the last tick's external current includes the objects
                    // sending every tick
                    double ext[2];
                    if ((rc = download(e, hc->dev.o2dExt + 16 * (int64_t)h, 2, ext))) return rc;
                    if (hc->stepsDone > 0) {
                        eg = ext[0];
                        ege = ext[1];
                    }
                }
                Im += ege - eg * V[h];
                out[q] = Im;
                break;
            }
            case HSG_F_INJECT: {
                double v = 0.0;
                for (size_t k = 0; k < hc->injIdx.size(); ++k) if (hc->injIdx[k] == h) v = hc->injBasal[k];
                out[q] = v;
                break;
            }
            case HSG_F_INITVM: out[q] = s.initVm[h]; break;
            case HSG_F_CM: out[q] = s.Cm[h]; break;
            case HSG_F_EM: out[q] = s.Em[h]; break;
            case HSG_F_RM: out[q] = s.Rm[h]; break;
            case HSG_F_RA: out[q] = s.Ra[h]; break;
            }
        } else if (field >= 20 && field <= 28) {
            if (i < 0 || i >= (int)s.chanCompt.size() || s.chanCompt[i] < 0)
                return fail(HSG_E_ARG, "unknown channel index");
            const int h = s.chanCompt[i];
            const size_t o = (size_t)s.chanSlot[i] * nc + h;
            auto gt = [&](double v) { return e->fp32 ? (double)(float)v : v; };
            auto stateAt = [&](int which) -> double {
                int js = s.chanState[i];
                if (which >= 1 && s.chanXp[i] > 0.0) ++js;
                if (which >= 2 && s.chanYp[i] > 0.0) ++js;
                return state[(size_t)js * nc + h];
            };
            switch (field) {
            case HSG_F_GBAR: out[q] = gt(s.Gbar[o]); break;
            case HSG_F_EK: out[q] = gt(s.Ek[o]); break;
            case HSG_F_GK:
            case HSG_F_IK: {
                if ((rc = needState())) return rc;
                auto ov = hc->gkOverride.find(i);
                double gk = ov != hc->gkOverride.end() ? ov->second
                          : e->fp32 ? chan_gk<float>(*hc, i, state) : chan_gk<double>(*hc, i, state);
                if (field == HSG_F_GK) out[q] = gk;
                else { if ((rc = needV())) return rc; out[q] = (gt(s.Ek[o]) - V[h]) * gk; }
                break;
            }
            case HSG_F_X: if ((rc = needState())) return rc; out[q] = s.chanXp[i] == 0.0 ? 0.0 : stateAt(0); break;
            case HSG_F_Y: if ((rc = needState())) return rc; out[q] = s.chanYp[i] == 0.0 ? 0.0 : stateAt(1); break;
            case HSG_F_Z: if ((rc = needState())) return rc; out[q] = s.chanZp[i] == 0.0 ? 0.0 : stateAt(2); break;
            case HSG_F_MODULATION: out[q] = s.mod[o]; break;
            case HSG_F_INSTANT: {
                int lt = s.chTypeLocal[o];
                out[q] = lt >= 0 ? (double)s.localTypes[lt].instant : 0.0;
                break;
            }
            default: return fail(HSG_E_ARG, "unknown field");
            }
        } else if (field >= 40 && field <= 43) {
            if (i < 0 || i >= (int)s.poolCompt.size() || s.poolCompt[i] < 0)
                return fail(HSG_E_ARG, "unknown pool index");
            const size_t o = (size_t)s.poolSlot[i] * nc + s.poolCompt[i];
            switch (field) {
            case HSG_F_CA:
                if (pool.empty()) {
                    pool.resize((size_t)s.poolSlots * nc);
                    if ((rc = download(e, hc->dev.oPoolCa, pool.size(), pool.data()))) return rc;
                }
                out[q] = pool[o];
                break;
            case HSG_F_CABASAL: out[q] = s.pBasal[o]; break;
            case HSG_F_CACEIL: out[q] = s.pCeil[o]; break;
            case HSG_F_CAFLOOR: out[q] = s.pFloor[o]; break;
            }
        } else if (field >= HSG_F_2D_GBAR && field <= HSG_F_2D_MODULATION) {
            // the HHChannel2D object's own fields
            if (i < 0 || i >= (int)s.ch2dOfDesc.size() || s.ch2dOfDesc[i] < 0)
                return fail(HSG_E_ARG, "unknown HHChannel2D index");
            const int k = s.ch2dOfDesc[i];
            const HostChan2D& hcd = s.ch2d[k];
            switch (field) {
            case HSG_F_2D_GBAR: out[q] = hcd.Gbar; break;
            case HSG_F_2D_EK: out[q] = hcd.Ek; break;
            case HSG_F_2D_MODULATION: out[q] = hcd.mod; break;
            case HSG_F_2D_GK:
                if ((rc = download(e, hc->dev.o2dGk + 8 * (int64_t)k, 1, &out[q]))) return rc;
                break;
            case HSG_F_2D_X: case HSG_F_2D_Y: case HSG_F_2D_Z:
                if ((rc = download(e, hc->dev.o2dState + 8 * (3 * (int64_t)k + (field - HSG_F_2D_X)), 1, &out[q])))
                    return rc;
                break;
            default: return fail(HSG_E_ARG, "unknown field");
            }
        } else if (field >= HSG_F_SYN_GBAR && field <= HSG_F_SYN_Y) {
            if (i < 0 || i >= (int)s.synOfDesc.size()) return fail(HSG_E_ARG, "unknown SynChan index");
            const int k = s.synOfDesc[i];
            const HostSyn& hs = s.syn[k];
            const int64_t rec = hc->dev.oSyn + (int64_t)sizeof(SynDev) * k;
            switch (field) {
            case HSG_F_SYN_GBAR: out[q] = hs.Gbar; break;
            case HSG_F_SYN_EK: out[q] = hs.Ek; break;
            case HSG_F_SYN_TAU1: out[q] = hs.tau1; break;
            case HSG_F_SYN_TAU2: out[q] = hs.tau2; break;
            case HSG_F_SYN_MODULATION: out[q] = hs.mod; break;
            case HSG_F_SYN_GK:
                if ((rc = download(e, rec + offsetof(SynDev, Gk), 1, &out[q]))) return rc;
                break;
            case HSG_F_SYN_X:
                if ((rc = download(e, rec + offsetof(SynDev, X), 1, &out[q]))) return rc;
                break;
            case HSG_F_SYN_Y:
                if ((rc = download(e, rec + offsetof(SynDev, Y), 1, &out[q]))) return rc;
                break;
            default: return fail(HSG_E_ARG, "unknown field");
            }
        } else if (field == HSG_F_SYNAPSE_WEIGHT || field == HSG_F_SYNAPSE_DELAY) {
            if (i < 0 || i >= (int)s.synW.size()) return fail(HSG_E_ARG, "unknown synapse index");
            out[q] = field == HSG_F_SYNAPSE_WEIGHT ? s.synW[i] : s.synDelay[i];
        } else if (field >= HSG_F_LEAK_GBAR && field <= HSG_F_LEAK_MODULATION) {
            if (i < 0 || i >= (int)s.leakOfDesc.size() || s.leakOfDesc[i] < 0)
                return fail(HSG_E_ARG, "unknown Leakage index");
            const HostLeak& l = s.leaks[s.leakOfDesc[i]];
            out[q] = field == HSG_F_LEAK_GBAR ? l.Gbar : field == HSG_F_LEAK_EK ? l.Ek : l.mod;
        } else if (field == HSG_F_VCLAMP_COMMAND || field == HSG_F_VCLAMP_CURRENT) {
            if (i < 0 || i >= (int)s.vclamps.size()) return fail(HSG_E_ARG, "unknown VClamp index");
            const int64_t rec = hc->dev.oVClamp + (int64_t)sizeof(VClampDev) * i;
            const int64_t off = field == HSG_F_VCLAMP_COMMAND ? offsetof(VClampDev, cmdIn) : offsetof(VClampDev, current);
            if ((rc = download(e, rec + off, 1, &out[q]))) return rc;
        } else if (field == HSG_F_PULSE_OUTPUT || field == HSG_F_PULSE_INPUT) {
            if (i < 0 || i >= (int)s.pulses.size()) return fail(HSG_E_ARG, "unknown PulseGen index");
            if (field == HSG_F_PULSE_OUTPUT && s.pulses[i].entry < 0) {
                // a free-running input source the device never steps: its
                // output at the last tick (its baseLevel before the first)
                int64_t steps = 0;
                if ((rc = download(e, hc->dev.oSteps, 1, &steps))) return rc;
                out[q] = steps == 0 ? s.pulses[i].base
                                    : pulse_free_output(s.pulses[i], e->dt * (double)(unsigned long long)steps);
                continue;
            }
            const int64_t rec = hc->dev.oPulse + (int64_t)sizeof(PulseDev) * i;
            const int64_t off = field == HSG_F_PULSE_OUTPUT ? offsetof(PulseDev, out) : offsetof(PulseDev, input);
            if ((rc = download(e, rec + off, 1, &out[q]))) return rc;
        } else {
            return fail(HSG_E_ARG, "unknown field");
        }
    }
    return 0;
}

int hsg_set_field(hsg_engine* e, int32_t cell, int32_t field, const int32_t* idx, int32_t n, const double* vals) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    if (n < 0 || (n > 0 && (!idx || !vals))) return fail(HSG_E_ARG, "bad index list");
    hipSetDevice(e->device);
    CellSetup& s = hc->s;
    const int nc = s.n;
    int rc;
    for (int q = 0; q < n; ++q) {
        const int i = idx[q];
        const double x = vals[q];
        if (field >= 1 && field <= 8) {
            if (i < 0 || i >= (int)s.descToHines.size() || s.descToHines[i] < 0)
                return fail(HSG_E_ARG, "unknown compartment index");
            const int h = s.descToHines[i];
            switch (field) {
            case HSG_F_VM: if ((rc = put1<double>(e, hc->dev.oV, h, x))) return rc; break;
            case HSG_F_IM: return fail(HSG_E_ARG, "Im is read-only");
            case HSG_F_INJECT: {
                bool found = false;
                for (size_t k = 0; k < hc->injIdx.size(); ++k)
                    if (hc->injIdx[k] == h) { hc->injBasal[k] = x; found = true; }
                if (!found) { hc->injIdx.push_back(h); hc->injBasal.push_back(x); hc->injVar.push_back(0.0); }
                hc->injDirty = true;
                break;
            }
            case HSG_F_INITVM:
                s.initVm[h] = x;
                if ((rc = put1<double>(e, hc->dev.oInitVm, h, x))) return rc;
                break;
            case HSG_F_CM:
                s.Cm[h] = x;
                s.CmByDt[h] = 2.0 * x / hc->hsolveDt;
                if ((rc = put1<double>(e, hc->dev.oCmByDt, h, s.CmByDt[h]))) return rc;
                break;
            case HSG_F_EM:
                s.Em[h] = x;
                s.EmByRm[h] = x / s.Rm[h];
                if ((rc = put1<double>(e, hc->dev.oEmByRm, h, s.EmByRm[h]))) return rc;
                break;
            case HSG_F_RM:
                s.Rm[h] = x;
                s.EmByRm[h] = s.Em[h] / x;
                if ((rc = put1<double>(e, hc->dev.oEmByRm, h, s.EmByRm[h]))) return rc;
                break;
            case HSG_F_RA: s.Ra[h] = x; break;   // no matrix rebuild, as the reference
            }
        } else if (field >= 20 && field <= 28) {
            if (i < 0 || i >= (int)s.chanCompt.size() || s.chanCompt[i] < 0)
                return fail(HSG_E_ARG, "unknown channel index");
            const int h = s.chanCompt[i];
            const size_t o = (size_t)s.chanSlot[i] * nc + h;
            auto stateIdx = [&](int which) -> size_t {
                int js = s.chanState[i];
                if (which >= 1 && s.chanXp[i] > 0.0) ++js;
                if (which >= 2 && s.chanYp[i] > 0.0) ++js;
                return (size_t)js * nc + h;
            };
            if ((field == HSG_F_GBAR || field == HSG_F_X || field == HSG_F_Y || field == HSG_F_Z ||
                 field == HSG_F_MODULATION) && !hc->gkOverride.count(i)) {
                std::vector<double> st;
                st.resize((size_t)s.gateSlots * nc);
                if (s.gateSlots > 0 && (rc = download(e, hc->dev.oState, st.size(), st.data()))) return rc;
                hc->gkOverride[i] = e->fp32 ? chan_gk<float>(*hc, i, st) : chan_gk<double>(*hc, i, st);
            }
            switch (field) {
            case HSG_F_GBAR: s.Gbar[o] = x; if ((rc = put_gt(e, hc->dev.oGbar, o, x))) return rc; break;
            case HSG_F_EK: s.Ek[o] = x; if ((rc = put_gt(e, hc->dev.oEk, o, x))) return rc; break;
            case HSG_F_GK: case HSG_F_IK: return fail(HSG_E_UNSUPPORTED, "Gk/Ik are recomputed every step");
            case HSG_F_X: if (s.chanXp[i] != 0.0 && (rc = put1<double>(e, hc->dev.oState, stateIdx(0), x))) return rc; break;
            case HSG_F_Y: if (s.chanYp[i] != 0.0 && (rc = put1<double>(e, hc->dev.oState, stateIdx(1), x))) return rc; break;
            case HSG_F_Z: if (s.chanZp[i] != 0.0 && (rc = put1<double>(e, hc->dev.oState, stateIdx(2), x))) return rc; break;
            case HSG_F_MODULATION:
                if (x > 0.0) {   // HSolve::setHHmodulation ignores non-positive values
                    s.mod[o] = x;
                    if ((rc = put1<double>(e, hc->dev.oMod, o, x))) return rc;
                    if (!hc->dev.hasModulation && x != 1.0) { hc->dev.hasModulation = 1; e->tablesDirty = true; }
                }
                break;
            case HSG_F_INSTANT: {
                int lt = s.chTypeLocal[o];
                HostChanType t = s.localTypes[lt];
                t.instant = (int)x;
                int nl = -1;
                for (size_t k = 0; k < s.localTypes.size(); ++k) {
                    const HostChanType& u = s.localTypes[k];
                    if (u.Xpower == t.Xpower && u.Ypower == t.Ypower && u.Zpower == t.Zpower && u.instant == t.instant &&
                        u.colX == t.colX && u.colY == t.colY && u.colZ == t.colZ && u.caDepLocal == t.caDepLocal &&
                        u.caTgtLocal == t.caTgtLocal) nl = (int)k;
                }
                if (nl < 0) {
                    nl = (int)s.localTypes.size();
                    s.localTypes.push_back(t);
                    hc->typeIds.push_back(type_id(e, t, hc->tableSet));
                }
                s.chTypeLocal[o] = nl;
                if ((rc = put1<int32_t>(e, hc->dev.oChType, o, hc->typeIds[nl]))) return rc;
                if (hc->dev.uniform) {
                    hc->dev.uniform = 0;
                    hc->dev.fastOk = 0;
                }
                e->tablesDirty = true;
                e->planDirty = true;
                break;
            }
            default: return fail(HSG_E_ARG, "unknown field");
            }
        } else if (field >= 40 && field <= 43) {
            if (i < 0 || i >= (int)s.poolCompt.size() || s.poolCompt[i] < 0)
                return fail(HSG_E_ARG, "unknown pool index");
            const size_t o = (size_t)s.poolSlot[i] * nc + s.poolCompt[i];
            switch (field) {
            case HSG_F_CA:
            case HSG_F_CABASAL:   // ZombieCaConc::vSetCaBasal calls setCa (hsolve/ZombieCaConc.cpp:62-65)
                if ((rc = put1<double>(e, hc->dev.oPoolCa, o, x))) return rc;
                s.pC[o] = x - s.pBasal[o];
                if ((rc = put1<double>(e, hc->dev.oPoolC, o, s.pC[o]))) return rc;
                break;
            case HSG_F_CACEIL: s.pCeil[o] = x; if ((rc = put1<double>(e, hc->dev.oPoolCeil, o, x))) return rc; break;
            case HSG_F_CAFLOOR: s.pFloor[o] = x; if ((rc = put1<double>(e, hc->dev.oPoolFloor, o, x))) return rc; break;
            }
        } else if (field >= HSG_F_2D_GBAR && field <= HSG_F_2D_MODULATION) {
            if (i < 0 || i >= (int)s.ch2dOfDesc.size() || s.ch2dOfDesc[i] < 0)
                return fail(HSG_E_ARG, "unknown HHChannel2D index");
            const int k = s.ch2dOfDesc[i];
            HostChan2D& hcd = s.ch2d[k];
            const int64_t rec = hc->dev.o2dChan + (int64_t)sizeof(Chan2DDev) * k;
            switch (field) {
            case HSG_F_2D_GBAR:
                hcd.Gbar = x;
                if ((rc = put1<double>(e, rec + offsetof(Chan2DDev, Gbar), 0, x))) return rc;
                break;
            case HSG_F_2D_EK:
                hcd.Ek = x;
                if ((rc = put1<double>(e, rec + offsetof(Chan2DDev, Ek), 0, x))) return rc;
                break;
            case HSG_F_2D_MODULATION:
                hcd.mod = x;
                if ((rc = put1<double>(e, rec + offsetof(Chan2DDev, mod), 0, x))) return rc;
                break;
            case HSG_F_2D_GK: return fail(HSG_E_UNSUPPORTED, "Gk is recomputed every step");
            case HSG_F_2D_X: case HSG_F_2D_Y: case HSG_F_2D_Z:
                if ((rc = put1<double>(e, hc->dev.o2dState, 3 * (size_t)k + (field - HSG_F_2D_X), x))) return rc;
                break;
            default: return fail(HSG_E_ARG, "unknown field");
            }
        } else if (field >= HSG_F_SYN_GBAR && field <= HSG_F_SYN_Y) {
            if (i < 0 || i >= (int)s.synOfDesc.size()) return fail(HSG_E_ARG, "unknown SynChan index");
            const int k = s.synOfDesc[i];
            HostSyn& hs = s.syn[k];
            const int64_t rec = hc->dev.oSyn + (int64_t)sizeof(SynDev) * k;
            switch (field) {
            case HSG_F_SYN_EK:
                hs.Ek = x;
                if ((rc = put1<double>(e, rec + offsetof(SynDev, Ek), 0, x))) return rc;
                break;
            case HSG_F_SYN_MODULATION:
                hs.mod = x;
                if ((rc = put1<double>(e, rec + offsetof(SynDev, mod), 0, x))) return rc;
                break;
            case HSG_F_SYN_GBAR: case HSG_F_SYN_TAU1: case HSG_F_SYN_TAU2: {
                if (field == HSG_F_SYN_TAU1 && !(x > 0.0)) return fail(HSG_E_ARG, "SynChan: tau1 must be > 0");
                if (field == HSG_F_SYN_TAU2 && !(x >= 0.0)) return fail(HSG_E_ARG, "SynChan: tau2 must be >= 0");
                (field == HSG_F_SYN_GBAR ? hs.Gbar : field == HSG_F_SYN_TAU1 ? hs.tau1 : hs.tau2) = x;
                // the setters recompute the constants at the object's dt_
                SynConst& c = hc->synConst[k];
                c = syn_constants(hs, c.dt);
                const double v[5] = {c.xconst1, c.xconst2, c.yconst1, c.yconst2, c.norm};
                HIPCHK(hipMemcpyAsync(e->arena.base + rec + offsetof(SynDev, xconst1), v, sizeof v,
                                      hipMemcpyHostToDevice, e->stream));
                HIPCHK(hipStreamSynchronize(e->stream));
                break;
            }
            default: return fail(HSG_E_ARG, "read-only SynChan field");
            }
        } else if (field == HSG_F_SYNAPSE_WEIGHT || field == HSG_F_SYNAPSE_DELAY) {
            if (i < 0 || i >= (int)s.synW.size()) return fail(HSG_E_ARG, "unknown synapse index");
            if (!std::isfinite(x) || (field == HSG_F_SYNAPSE_DELAY && !(x >= 0.0)))
                return fail(HSG_E_ARG, "synapse weight must be finite, delay finite and >= 0");
            (field == HSG_F_SYNAPSE_WEIGHT ? s.synW[i] : s.synDelay[i]) = x;
        } else if (field >= HSG_F_LEAK_GBAR && field <= HSG_F_LEAK_MODULATION) {
            if (i < 0 || i >= (int)s.leakOfDesc.size() || s.leakOfDesc[i] < 0)
                return fail(HSG_E_ARG, "unknown Leakage index");
            const int k = s.leakOfDesc[i];
            HostLeak& l = s.leaks[k];
            (field == HSG_F_LEAK_GBAR ? l.Gbar : field == HSG_F_LEAK_EK ? l.Ek : l.mod) = x;
            const LeakDev rec = leak_records(s)[k];
            HIPCHK(hipMemcpyAsync(e->arena.base + hc->dev.oLeak + (int64_t)sizeof(LeakDev) * k, &rec, sizeof rec,
                                  hipMemcpyHostToDevice, e->stream));
            HIPCHK(hipStreamSynchronize(e->stream));
        } else if (field == HSG_F_VCLAMP_COMMAND) {
            // VClamp::setCommand: cmdIn_
            if (i < 0 || i >= (int)s.vclamps.size()) return fail(HSG_E_ARG, "unknown VClamp index");
            const int64_t rec = hc->dev.oVClamp + (int64_t)sizeof(VClampDev) * i;
            if ((rc = put1<double>(e, rec + offsetof(VClampDev, cmdIn), 0, x))) return rc;
        } else if (field == HSG_F_PULSE_INPUT) {
            // PulseGen::input (device/PulseGen.cpp:355-358), where no
            // PulseGen's output drives it
            if (i < 0 || i >= (int)s.pulses.size()) return fail(HSG_E_ARG, "unknown PulseGen index");
            if (s.pulses[i].src >= 0) return fail(HSG_E_ARG, "this PulseGen's input comes from another PulseGen");
            const int64_t rec = hc->dev.oPulse + (int64_t)sizeof(PulseDev) * i;
            if ((rc = put1<double>(e, rec + offsetof(PulseDev, input), 0, x))) return rc;
        } else {
            return fail(HSG_E_ARG, "unknown field");
        }
    }
    return 0;
}

int hsg_add_inject(hsg_engine* e, int32_t cell, int32_t compartment, double value) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    const CellSetup& s = hc->s;
    if (compartment < 0 || compartment >= (int)s.descToHines.size() || s.descToHines[compartment] < 0)
        return fail(HSG_E_ARG, "unknown compartment index");
    const int h = s.descToHines[compartment];
    for (size_t k = 0; k < hc->injIdx.size(); ++k)
        if (hc->injIdx[k] == h) { hc->injVar[k] += value; hc->injDirty = true; return 0; }
    hc->injIdx.push_back(h);
    hc->injBasal.push_back(0.0);
    hc->injVar.push_back(0.0 + value);
    hc->injDirty = true;
    return 0;
}

int hsg_set_pulsegen(hsg_engine* e, int32_t cell, int32_t pulsegen, const hsg_pulsegen_desc* desc) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    CellSetup& s = hc->s;
    if (pulsegen < 0 || pulsegen >= (int)s.pulses.size() || !desc) return fail(HSG_E_ARG, "no such PulseGen");
    if (desc->n_pulses < 1 || desc->n_pulses > HSG_MAX_PULSES)
        return fail(HSG_E_ARG, "PulseGen: 1 to HSG_MAX_PULSES pulses");
    HostPulse& p = s.pulses[pulsegen];
    const int compt = desc->compartment >= 0 && desc->compartment < (int)s.descToHines.size()
                          ? s.descToHines[desc->compartment] : -2;
    if (desc->compartment < -1 || (desc->compartment == -1 ? p.compt != -1 : compt != p.compt))
        return fail(HSG_E_ARG, "a PulseGen keeps the compartment it was added with");
    if (desc->trigMode < 0 || desc->trigMode > 2) return fail(HSG_E_ARG, "PulseGen: trigMode 0, 1 or 2");
    const int src = desc->trigMode != 0 ? desc->input_src : -1;
    if (src < -1 || src >= (int)s.pulses.size() || src == pulsegen)
        return fail(HSG_E_ARG, "PulseGen: input_src is -1 or another PulseGen of the cell");
    // a source must stay one (this one, if others read it, and its own)
    HostPulse next = p;
    next.n = desc->n_pulses;
    next.trigMode = desc->trigMode;
    next.base = desc->baseLevel;
    for (int j = 0; j < HSG_MAX_PULSES; ++j) {
        next.delay[j] = j < next.n ? desc->delay[j] : 0.0;
        next.width[j] = j < next.n ? desc->width[j] : 0.0;
    }
    for (size_t k = 0; k < s.pulses.size(); ++k)
        if ((int)k != pulsegen && s.pulses[k].src == pulsegen && !pulse_source_ok(next))
            return fail(HSG_E_UNSUPPORTED, "PulseGen: an input source must run free with pulses that cover its period");
    if (src >= 0 && !pulse_source_ok(s.pulses[src]))
        return fail(HSG_E_UNSUPPORTED, "PulseGen: an input source must run free with pulses that cover its period");
    p.trigMode = desc->trigMode;
    p.src = src;
    p.n = desc->n_pulses;
    p.base = desc->baseLevel;
    for (int j = 0; j < HSG_MAX_PULSES; ++j) {
        p.delay[j] = j < p.n ? desc->delay[j] : 0.0;
        p.width[j] = j < p.n ? desc->width[j] : 0.0;
        p.level[j] = j < p.n ? desc->level[j] : 0.0;
    }
    hipSetDevice(e->device);
    // everything but the state the device holds (output, input, trigger
    // history)
    PulseDev q{};
    const int64_t off = hc->dev.oPulse + (int64_t)sizeof(PulseDev) * pulsegen;
    HIPCHK(hipMemcpyAsync(&q, e->arena.base + off, sizeof q, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    q.n = p.n;
    q.trigMode = p.trigMode;
    q.src = p.src;
    q.srcBefore = p.src >= 0 && p.src < pulsegen;
    q.base = p.base;
    for (int j = 0; j < HSG_MAX_PULSES; ++j) {
        q.delay[j] = p.delay[j];
        q.width[j] = p.width[j];
        q.level[j] = p.level[j];
    }
    HIPCHK(hipMemcpyAsync(e->arena.base + off, &q, sizeof q, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

int hsg_add_spike(hsg_engine* e, int32_t cell, int32_t synchan, int32_t synapse, double time) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    const CellSetup& s = hc->s;
    if (synchan < 0 || synchan >= (int)s.synOfDesc.size()) return fail(HSG_E_ARG, "no such SynChan");
    const int q = s.synOfDesc[synchan];
    if (synapse < 0 || synapse >= s.syn[q].nSyn) return fail(HSG_E_ARG, "no such synapse");
    if (!std::isfinite(time)) return fail(HSG_E_ARG, "spike time must be finite");
    const int j = s.syn[q].firstSyn + synapse;
    queue_event(hc->synQ[q], time + s.synDelay[j], s.synW[j]);
    return 0;
}

int hsg_connect_spike(hsg_engine* e, int32_t src_cell, int32_t spikegen, int32_t dst_cell, int32_t synchan,
                      int32_t synapse) {
    HostCell* src = get_cell(e, src_cell);
    HostCell* dst = get_cell(e, dst_cell);
    if (!src || !dst) return fail(HSG_E_ARG, "no such cell");
    bool found = false;
    for (int g : src->s.spkDesc) found = found || g == spikegen;
    if (!found) return fail(HSG_E_ARG, "no such SpikeGen in the source cell");
    const CellSetup& s = dst->s;
    if (synchan < 0 || synchan >= (int)s.synOfDesc.size()) return fail(HSG_E_ARG, "no such SynChan");
    if (synapse < 0 || synapse >= s.syn[s.synOfDesc[synchan]].nSyn) return fail(HSG_E_ARG, "no such synapse");
    e->conns[{src_cell, spikegen}].push_back(hsg_engine::SpikeTarget{dst_cell, synchan, synapse});
    return 0;
}

int hsg_add_gkek(hsg_engine* e, int32_t cell, int32_t compartment, double Gk, double Ek) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    const CellSetup& s = hc->s;
    if (compartment < 0 || compartment >= (int)s.descToHines.size() || s.descToHines[compartment] < 0)
        return fail(HSG_E_ARG, "unknown compartment index");
    auto& v = hc->ext[s.descToHines[compartment]];
    v.first += Gk;
    v.second += Gk * Ek;
    return 0;
}

int hsg_add_conc(hsg_engine* e, int32_t cell, int32_t channel, double conc) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    const CellSetup& s = hc->s;
    if (channel < 0 || channel >= (int)s.chanCompt.size() || s.chanCompt[channel] < 0)
        return fail(HSG_E_ARG, "unknown channel index");
    hipSetDevice(e->device);
    return put1<double>(e, hc->dev.oExtCa, (size_t)s.chanSlot[channel] * s.n + s.chanCompt[channel], conc);
}

int hsg_add_ica(hsg_engine* e, int32_t cell, int32_t pool, double iCa) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    const CellSetup& s = hc->s;
    if (pool < 0 || pool >= (int)s.poolCompt.size() || s.poolCompt[pool] < 0)
        return fail(HSG_E_ARG, "unknown pool index");
    hc->ica[pool] += iCa;
    return 0;
}

int hsg_read_spikes(hsg_engine* e, hsg_spike* out, int32_t cap, int32_t* n_out, int64_t* dropped) {
    if (!e || !n_out || (cap > 0 && !out)) return fail(HSG_E_ARG, "null argument");
    hipSetDevice(e->device);
    unsigned long long cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, e->dSpikeCount, sizeof cnt, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    const unsigned long long have = std::min<unsigned long long>(cnt, kSpikeCap);
    std::vector<SpikeRec> recs(have);
    if (have)
        HIPCHK(hipMemcpyAsync(recs.data(), e->dSpikes, sizeof(SpikeRec) * have, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemsetAsync(e->dSpikeCount, 0, sizeof(unsigned long long), e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->spikesDropped += (int64_t)(cnt - have);
    recs.insert(recs.end(), e->spikeBuf.begin(), e->spikeBuf.end());
    e->spikeBuf.clear();
    std::sort(recs.begin(), recs.end(), [](const SpikeRec& a, const SpikeRec& b) {
        if (a.step != b.step) return a.step < b.step;
        if (a.cell != b.cell) return a.cell < b.cell;
        return a.gen < b.gen;
    });
    int32_t w = 0;
    for (const SpikeRec& r : recs) {
        if (w >= cap) { e->spikesDropped++; continue; }
        HostCell* hc = get_cell(e, r.cell);
        out[w].cell = r.cell;
        out[w].spikegen = hc && r.gen < (int)hc->s.spkDesc.size() ? hc->s.spkDesc[r.gen] : r.gen;
        out[w].step = r.step;
        ++w;
    }
    *n_out = w;
    if (dropped) { *dropped = e->spikesDropped; e->spikesDropped = 0; }
    return 0;
}

int hsg_set_probes(hsg_engine* e, int32_t cell, const int32_t* compartments, int32_t n, int64_t capacity) {
    return hsg_set_probes_every(e, cell, compartments, n, capacity, 1);
}

// A probe set: the arena offsets of the doubles sampled every `every` steps.
static int set_probe_offsets(hsg_engine* e, HostCell* hc, const std::vector<int64_t>& offs, int64_t capacity,
                             int32_t every) {
    const int n = (int)offs.size();
    if (hc->probe.off >= 0) { e->arena.release(hc->probe.off, hc->probe.bytes); hc->probe.off = -1; }
    hc->dev.nProbe = 0;
    hc->dev.probeCap = 0;
    hc->dev.probeStride = every;
    e->tablesDirty = true;
    if (n == 0 || capacity == 0) return 0;
    int64_t bytes = align_up(8 * (int64_t)n) + align_up(8 * (int64_t)n * capacity) + kAlign;
    int64_t off;
    int rc = arena_alloc(e, bytes, &off);
    if (rc) return rc;
    hc->probe.off = off;
    hc->probe.bytes = bytes;
    hc->dev.oProbeOff = off;
    hc->dev.oProbeTrace = off + align_up(8 * (int64_t)n);
    hc->dev.oProbeBase = hc->dev.oProbeTrace + align_up(8 * (int64_t)n * capacity);
    if ((rc = upload(e, hc->dev.oProbeOff, offs))) return rc;
    if ((rc = put1<int64_t>(e, hc->dev.oProbeBase, 0, hc->stepsDone))) return rc;
    hc->dev.nProbe = n;
    hc->dev.probeCap = capacity;
    hc->probeBase = hc->stepsDone;
    hc->probeCompt.assign(n, 0);
    hc->probeCap = capacity;
    return 0;
}

int hsg_set_probes_every(hsg_engine* e, int32_t cell, const int32_t* compartments, int32_t n, int64_t capacity,
                         int32_t every) {
    if (n < 0 || (n > 0 && !compartments)) return fail(HSG_E_ARG, "bad probe list");
    std::vector<int32_t> fields(n > 0 ? n : 0, HSG_F_VM);
    return hsg_set_probe_fields(e, cell, fields.data(), compartments, n, capacity, every);
}

int hsg_set_probe_fields(hsg_engine* e, int32_t cell, const int32_t* fields, const int32_t* idx, int32_t n,
                         int64_t capacity, int32_t every) {
    HostCell* hc = get_cell(e, cell);
    if (!hc) return fail(HSG_E_ARG, "no such cell");
    if (n < 0 || capacity < 0 || (n > 0 && (!fields || !idx))) return fail(HSG_E_ARG, "bad probe list");
    if (every < 1) return fail(HSG_E_ARG, "probe interval must be at least one step");
    hipSetDevice(e->device);
    const CellSetup& s = hc->s;
    const CellDev& d = hc->dev;
    std::vector<int64_t> offs(n);
    for (int k = 0; k < n; ++k) {
        const int i = idx[k];
        switch (fields[k]) {
        case HSG_F_VM:
            if (i < 0 || i >= (int)s.descToHines.size() || s.descToHines[i] < 0)
                return fail(HSG_E_ARG, "unknown compartment index");
            offs[k] = d.oV + 8 * (int64_t)s.descToHines[i];
            break;
        case HSG_F_CA:
            if (i < 0 || i >= (int)s.poolCompt.size() || s.poolCompt[i] < 0)
                return fail(HSG_E_ARG, "unknown pool index");
            offs[k] = d.oPoolCa + 8 * ((int64_t)s.poolSlot[i] * s.n + s.poolCompt[i]);
            break;
        case HSG_F_VCLAMP_CURRENT:
            if (i < 0 || i >= (int)s.vclamps.size()) return fail(HSG_E_ARG, "unknown VClamp index");
            offs[k] = d.oVClamp + (int64_t)sizeof(VClampDev) * i + (int64_t)offsetof(VClampDev, current);
            break;
        case HSG_F_SYN_GK:
            if (i < 0 || i >= (int)s.synOfDesc.size()) return fail(HSG_E_ARG, "unknown SynChan index");
            offs[k] = d.oSyn + (int64_t)sizeof(SynDev) * s.synOfDesc[i] + (int64_t)offsetof(SynDev, Gk);
            break;
        default:
            return fail(HSG_E_ARG, "a probe samples Vm, Ca, a VClamp's current or a SynChan's Gk");
        }
    }
    return set_probe_offsets(e, hc, offs, capacity, every);
}

int hsg_read_probes(hsg_engine* e, int32_t cell, double* out, int64_t cap_samples, int64_t* n_samples) {
    HostCell* hc = get_cell(e, cell);
    if (!hc || !n_samples) return fail(HSG_E_ARG, "no such cell");
    *n_samples = 0;
    if (hc->dev.nProbe == 0) return 0;
    hipSetDevice(e->device);
    // sample k is the (k+1)-th multiple of the stride after the trace's
    // base step, written while below capacity
    const int64_t st = hc->dev.probeStride;
    const int64_t cnt = hc->stepsDone / st - hc->probeBase / st;
    int64_t k = std::min(std::min(cnt, hc->dev.probeCap), cap_samples);
    int rc;
    if (k > 0 && out && (rc = download(e, hc->dev.oProbeTrace, (size_t)(k * hc->dev.nProbe), out))) return rc;
    // the next trace starts here: only the device's base word changes, the
    // cell records and the launch plan stay as they are
    if (hc->probeBase != hc->stepsDone) {
        hc->probeBase = hc->stepsDone;
        if ((rc = put1<int64_t>(e, hc->dev.oProbeBase, 0, hc->stepsDone))) return rc;
    }
    *n_samples = std::max<int64_t>(0, k);
    return 0;
}

int hsg_cell_info_get(hsg_engine* e, int32_t cell, hsg_cell_info* out) {
    HostCell* hc = get_cell(e, cell);
    if (!hc || !out) return fail(HSG_E_ARG, "no such cell");
    const CellSetup& s = hc->s;
    out->n_compartments = s.n;
    int nch = 0;
    for (int c : s.chanCompt) nch += c >= 0;
    out->n_channels = nch;
    int ng = 0;
    for (size_t k = 0; k < s.chanCompt.size(); ++k)
        if (s.chanCompt[k] >= 0) ng += (s.chanXp[k] > 0) + (s.chanYp[k] > 0) + (s.chanZp[k] > 0);
    out->n_gates = ng;
    int np = 0;
    for (int c : s.poolCompt) np += c >= 0;
    out->n_pools = np;
    out->n_spikegens = (int)s.spkCompt.size();
    out->n_levels = (int)s.levelStart.size() - 1;
    out->max_group = s.maxGroup;
    out->device_bytes = hc->region.bytes + std::max<int64_t>(0, hc->inj.bytes) +
                        std::max<int64_t>(0, hc->probe.bytes) + std::max<int64_t>(0, hc->extBlk.bytes);
    out->steps_done = hc->stepsDone;
    out->vMin = s.tables.v.min;
    out->vMax = s.tables.v.max;
    out->vDiv = (int32_t)s.tables.v.nDivs;
    out->caMin = s.tables.ca.min;
    out->caMax = s.tables.ca.max;
    out->caDiv = (int32_t)s.tables.ca.nDivs;
    return 0;
}

int hsg_hines_order(hsg_engine* e, int32_t cell, int32_t* out) {
    HostCell* hc = get_cell(e, cell);
    if (!hc || !out) return fail(HSG_E_ARG, "no such cell");
    for (size_t i = 0; i < hc->s.descToHines.size(); ++i) out[i] = hc->s.descToHines[i];
    return 0;
}

int hsg_plan_info(hsg_engine* e, hsg_bin_info* out, int32_t cap, int32_t* n_out) {
    if (!e || !n_out || (cap > 0 && !out)) return fail(HSG_E_ARG, "null argument");
    hipSetDevice(e->device);
    int rc = upload_tables(e);   // the plan of the current live set
    if (rc) return rc;
    const auto& bins = e->plan.bins;
    *n_out = (int32_t)bins.size();
    for (int32_t b = 0; b < cap && b < (int32_t)bins.size(); ++b) {
        const hsg_engine::Bin& x = bins[b];
        hsg_bin_info& o = out[b];
        std::memset(&o, 0, sizeof o);
        o.cells = (int32_t)x.cells.size();
        o.threads = x.threads;
        o.solve_threads = x.split ? (x.msg ? x.threads : x.blocks ? 256 : x.solveThreads) : 0;
        o.lds_bytes = x.lds;
        o.split = x.split;
        o.msg = x.msg;
        o.fast_slots = x.split ? (x.fastSel == 0 ? 0 : x.chanMS) : fused_fast(x);
        o.slow_groups = x.slow;
        o.ext = x.hasExt;
        o.round_steps = x.roundSteps;
        o.compartments = x.sumN;
        o.round_ms = x.roundMs;
        const char* prec = e->fp32 ? "float" : "double";
        if (x.split && x.blocks)
            snprintf(o.kernel, sizeof o.kernel, "hsolve_block_fwd/top/bwd<%s>", prec);
        else if (x.split)
            snprintf(o.kernel, sizeof o.kernel, "%s<%s>", x.msg ? "hsolve_msg_kernel"
                     : x.solveThreads > 256 ? "hsolve_solve_wide_kernel" : "hsolve_solve_kernel", prec);
        else if (fused_fast(x) == 4)
            snprintf(o.kernel, sizeof o.kernel, "hsolve_step_ca_kernel<%s, %s>", prec, x.slow ? "true" : "false");
        else
            snprintf(o.kernel, sizeof o.kernel, "%s<%s%s>", x.slow ? "hsolve_step_slow_kernel"
                     : fused_fast(x) ? (x.lean && !x.hasExt ? "hsolve_step_lean_kernel" : "hsolve_step_fast_kernel")
                     : "hsolve_step_kernel", prec,
                     !x.slow && fused_fast(x) && x.lean && !x.hasExt ? (x.lean == 2 ? ", kSigHH" : ", 0") : "");
    }
    return 0;
}

int hsg_set_bin_shape(hsg_engine* e, int32_t bin, int32_t threads, int32_t fast_pass) {
    if (!e) return fail(HSG_E_ARG, "null engine");
    if (bin == -1) {
        e->shapes.clear();
        e->planDirty = true;
        return 0;
    }
    hipSetDevice(e->device);
    int rc = upload_tables(e);   // the plan the index refers to
    if (rc) return rc;
    if (bin < 0 || bin >= (int32_t)e->plan.bins.size()) return fail(HSG_E_ARG, "no such bin in the launch plan");
    if (threads != 0 && threads != 64 && threads != 128 && threads != 256)
        return fail(HSG_E_ARG, "threads per cell workgroup: 0, 64, 128 or 256");
    if (fast_pass < -1 || fast_pass > 1) return fail(HSG_E_ARG, "fast_pass: -1, 0 or 1");
    const hsg_engine::Bin& b = e->plan.bins[bin];
    if (threads != 0 && b.split && !b.msg && b.blocks)
        return fail(HSG_E_ARG, "the block-form kernels have a fixed workgroup size");
    const bool fastOk = b.split ? b.chanMS > 0 : ((b.chanMS == 2 && !b.slow) || (b.chanMS == 4 && b.sig == 2));
    if (fast_pass == 1 && !fastOk) return fail(HSG_E_ARG, "the fast channel pass does not apply to this bin's cells");
    hsg_engine::BinShape& sh = e->shapes[b.key];
    sh.threads = threads;
    sh.fast = fast_pass;
    e->planDirty = true;
    return 0;
}

int hsg_sync(hsg_engine* e) {
    if (!e) return fail(HSG_E_ARG, "null engine");
    hipSetDevice(e->device);
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

int hsg_last_step_timing(hsg_engine* e, int32_t* launches, double* total_ms, double* mean_ms) {
    if (!e) return fail(HSG_E_ARG, "null engine");
    if (launches) *launches = e->lastLaunches;
    if (total_ms) *total_ms = e->lastTotalMs;
    if (mean_ms) *mean_ms = e->lastLaunches ? e->lastTotalMs / e->lastLaunches : 0.0;
    return 0;
}

}  // extern "C"
