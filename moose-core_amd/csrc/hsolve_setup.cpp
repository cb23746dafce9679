// hsolve_setup.cpp — turns a cell described in MOOSE object terms into the
// Hines-ordered arrays and the level schedule of the GPU path.
//
// The numbering, matrix coefficients, table construction and channel/pool
// bookkeeping follow the reference's setup exactly (same arithmetic in the
// same order), because the FP64 GPU form is meant to reproduce the
// reference's numbers bit for bit:
//   walkTree ............ hsolve/HSolvePassive.cpp:45-98
//   initialize/storeTree  hsolve/HSolvePassive.cpp:100-205
//   makeJunctions/Matrix  hsolve/HinesMatrix.cpp:85-222
//   readHHChannels ...... hsolve/HSolveActiveSetup.cpp:162-233
//   readGates/Calcium ... hsolve/HSolveActiveSetup.cpp:235-337
//   createLookupTables .. hsolve/HSolveActiveSetup.cpp:339-526
//   readSynapses ........ hsolve/HSolveActiveSetup.cpp:538-578 (SpikeGens)
// The level schedule (which units may run concurrently) is this project's
// own; see schedule() below.
#include "hsolve_setup.h"

#include "hsolve_internal.h"

#include <algorithm>
#include <map>
#include <array>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <set>

namespace hsg {

namespace {

// Message lists per compartment, in message order.
struct Adjacency {
    std::vector<std::vector<int>> axial;    // children (axial targets)
    std::vector<std::vector<int>> raxial;   // parents (raxial targets)
};

Adjacency make_adjacency(const hsg_cell_desc* d) {
    Adjacency a;
    a.axial.resize(d->n_compartments);
    a.raxial.resize(d->n_compartments);
    for (int k = 0; k < d->n_axial; ++k) {
        int p = d->axial[2 * k], c = d->axial[2 * k + 1];
        a.axial[p].push_back(c);
        a.raxial[c].push_back(p);
    }
    return a;
}

// HSolveUtils::adjacent (hsolve/HSolveUtils.cpp:26-45) minus `exclude`.
void adjacent(const Adjacency& a, int c, int exclude, std::vector<int>& out) {
    for (int x : a.axial[c]) if (x != exclude) out.push_back(x);
    for (int x : a.raxial[c]) if (x != exclude) out.push_back(x);
}

// HSolvePassive::walkTree: find a terminal by following the first neighbour,
// then depth-first from it popping the last neighbour first; the visit order
// reversed is the Hines numbering.
int walk_tree(const hsg_cell_desc* d, const Adjacency& a, std::vector<int>& order, std::string& err) {
    const int n = d->n_compartments;
    int seed = d->seed;
    std::vector<int> adj;
    adjacent(a, seed, -2, adj);
    if (adj.size() > 1) {
        int guard = 0;
        while (!adj.empty()) {
            int prev = seed;
            seed = adj[0];
            adj.clear();
            adjacent(a, seed, prev, adj);
            if (++guard > n + 1) { err = "walkTree: the compartments form a loop"; return HSG_E_MODEL; }
        }
    }
    std::vector<std::vector<int>> stack;
    stack.push_back({seed});
    int above = -2;
    order.clear();
    while (!stack.empty()) {
        std::vector<int>& top = stack.back();
        if (top.empty()) {
            stack.pop_back();
            if (!stack.empty()) stack.back().pop_back();
            continue;
        }
        if (stack.size() > 1) above = stack[stack.size() - 2].back();
        int cur = top.back();
        if ((int)order.size() >= n) { err = "walkTree: the compartments form a loop"; return HSG_E_MODEL; }
        order.push_back(cur);
        std::vector<int> next;
        adjacent(a, cur, above, next);
        stack.push_back(std::move(next));
    }
    std::reverse(order.begin(), order.end());
    return 0;
}

// HHGate::lookupBoth with interpolation forced on (biophysics/HHGate.cpp:258-281)
void lookup_both(const hsg_gate_desc& g, double v, double& A, double& B) {
    const double invDx = (double)g.divs / (g.max - g.min);
    if (v <= g.min) { A = g.tableA[0]; B = g.tableB[0]; }
    else if (v >= g.max) { A = g.tableA[g.divs]; B = g.tableB[g.divs]; }
    else {
        unsigned idx = (unsigned)((v - g.min) * invDx);
        double frac = (v - g.min - idx / invDx) * invDx;
        A = g.tableA[idx] * (1 - frac) + g.tableA[idx + 1] * frac;
        B = g.tableB[idx] * (1 - frac) + g.tableB[idx + 1] * frac;
    }
}

// HSolveUtils::rates (hsolve/HSolveUtils.cpp:205-252): copy when the grids
// match, else resample.
void rates(const hsg_gate_desc& g, double mn, double mx, unsigned divs, std::vector<double>& A,
           std::vector<double>& B) {
    A.assign(divs + 1, 0.0);
    B.assign(divs + 1, 0.0);
    if (mn == g.min && mx == g.max && divs == g.divs) {
        std::copy(g.tableA, g.tableA + g.divs + 1, A.begin());
        std::copy(g.tableB, g.tableB + g.divs + 1, B.begin());
        return;
    }
    const double dx = (mx - mn) / divs;
    for (unsigned i = 0; i < divs + 1; ++i) lookup_both(g, mn + dx * i, A[i], B[i]);
}

// LookupTable ctor + addColumns (hsolve/RateLookup.cpp:15-55)
void make_table(HostTable& t, double mn, double mx, unsigned nDivs, unsigned nSpecies) {
    t.min = mn;
    t.max = mx;
    t.nDivs = nDivs;
    t.nSpecies = nSpecies;
    t.dx = (mx - mn) / nDivs;
    t.table.assign((size_t)t.nPts() * t.nColumns(), 0.0);
}
void add_columns(HostTable& t, int species, const std::vector<double>& A, const std::vector<double>& B) {
    const uint32_t nc = t.nColumns();
    for (uint32_t ig = 0; ig + 1 < t.nPts(); ++ig) {
        t.table[(size_t)ig * nc + 2 * species] = A[ig];
        t.table[(size_t)ig * nc + 2 * species + 1] = B[ig];
    }
    t.table[(size_t)(t.nPts() - 1) * nc + 2 * species] = A.back();
    t.table[(size_t)(t.nPts() - 1) * nc + 2 * species + 1] = B.back();
}

uint64_t fnv(const void* p, size_t n, uint64_t h) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
    return h;
}

}  // namespace

// ---------------------------------------------------------------------------
// Level schedule of the elimination.
//
// Every node but the last (the Hines root, n-1) is eliminated exactly once:
// by a linear unit (x -> x+1) or as a non-last member of one junction group.
// The only row updates a node receives from outside its own eliminating unit
// come from one "feeder": the linear unit of x-1, or the group in which the
// node is the last member.  Units therefore form a tree (a unit's parent is
// the unit that eliminates its target), and a level = height in that tree is
// a valid parallel schedule with disjoint writes per level.
//
// Order of updates: for a group member c_k (k >= 1) the reference subtracts
// the terms of its earlier siblings c_1..c_{k-1} first and the terms of its
// feeder afterwards (the feeder's nodes have higher Hines indices than
// c_{k-1}).  Such a feeder therefore parks its terms in a deferred slot that
// the group applies right before it pivots c_k.
// ---------------------------------------------------------------------------
static int schedule(CellSetup& s, const std::vector<int>& junctionGroup, std::string& err) {
    const int n = s.n;
    // eliminating unit of every node; feeder unit of every node
    std::vector<int> elimUnit(n, -1), feeder(n, -1);
    std::vector<Unit> units;
    for (int g = 0; g < (int)s.groups.size(); ++g) {
        Unit u{1, -1, g, 0};
        units.push_back(u);
    }
    for (int g = 0; g < (int)s.groups.size(); ++g) {
        const auto& m = s.groups[g].mem;
        for (size_t p = 0; p + 1 < m.size(); ++p) {
            if (elimUnit[m[p]] >= 0) { err = "junction structure: node eliminated twice"; return HSG_E_UNSUPPORTED; }
            elimUnit[m[p]] = g;
        }
    }
    for (int x = 0; x + 1 < n; ++x) {
        if (junctionGroup[x] >= 0) continue;
        Unit u{0, x, -1, 0};
        elimUnit[x] = (int)units.size();
        units.push_back(u);
    }
    for (int x = 0; x + 1 < n; ++x)
        if (elimUnit[x] < 0) { err = "junction structure: node never eliminated"; return HSG_E_UNSUPPORTED; }
    if (n >= 1 && elimUnit[n - 1] >= 0) {
        err = "junction structure: the last Hines node is not the root";
        return HSG_E_UNSUPPORTED;
    }
    // feeders
    for (int ui = 0; ui < (int)units.size(); ++ui) {
        const Unit& u = units[ui];
        int target = u.kind == 0 ? u.node + 1 : s.groups[u.group].mem.back();
        if (feeder[target] >= 0) {
            err = "junction structure: a row has two feeding units (not a walkTree order)";
            return HSG_E_UNSUPPORTED;
        }
        feeder[target] = ui;
    }
    // levels: process units by their largest eliminated node
    std::vector<int> key(units.size());
    for (size_t ui = 0; ui < units.size(); ++ui) {
        const Unit& u = units[ui];
        key[ui] = u.kind == 0 ? u.node : s.groups[u.group].mem[s.groups[u.group].mem.size() - 2];
    }
    std::vector<int> ord(units.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
    for (int ui : ord) {
        Unit& u = units[ui];
        int lv = 0;
        auto feed = [&](int node) {
            int f = feeder[node];
            if (f >= 0 && f != ui) {
                if (key[f] >= key[ui]) return false;
                lv = std::max(lv, units[f].level + 1);
            }
            return true;
        };
        bool ok = true;
        if (u.kind == 0) ok = feed(u.node);
        else {
            const auto& m = s.groups[u.group].mem;
            for (size_t p = 0; p + 1 < m.size(); ++p) ok = ok && feed(m[p]);
        }
        if (!ok) { err = "junction structure: feeder after its target"; return HSG_E_UNSUPPORTED; }
        u.level = lv;
    }
    // deferred terms (a group's members get consecutive slots, in member order)
    s.defer.assign(2 * (size_t)n, -1);
    for (int x = 0; x < n; ++x) s.defer[2 * x + 1] = 0;
    for (auto& g : s.groups) g.deferLast = -1;
    s.nTerm = 0;
    std::vector<int> termBase(s.groups.size(), 0);
    for (int g = 0; g < (int)s.groups.size(); ++g) {
        const auto& m = s.groups[g].mem;
        termBase[g] = s.nTerm;
        for (size_t p = 1; p + 1 < m.size(); ++p) {
            int f = feeder[m[p]];
            if (f < 0) continue;
            int cnt = units[f].kind == 0 ? 1 : (int)s.groups[units[f].group].mem.size() - 1;
            s.defer[2 * m[p]] = s.nTerm;
            s.defer[2 * m[p] + 1] = cnt;
            if (units[f].kind == 1) s.groups[units[f].group].deferLast = s.nTerm;
            s.nTerm += cnt;
        }
    }
    int nLevels = 0;
    for (const Unit& u : units) nLevels = std::max(nLevels, u.level + 1);
    s.levelStart.assign(nLevels + 1, 0);
    for (const Unit& u : units) s.levelStart[u.level + 1]++;
    for (int l = 0; l < nLevels; ++l) s.levelStart[l + 1] += s.levelStart[l];
    s.units.assign(units.size(), 0);
    std::vector<int> fill(s.levelStart.begin(), s.levelStart.end() - 1);
    for (int ui : ord) {
        const Unit& u = units[ui];
        s.units[fill[u.level]++] = u.kind == 0 ? u.node : -(u.group + 1);
    }
    // descriptors in the same (level) order; within a level, fast units first
    s.udesc.assign((size_t)kDescInts * s.units.size(), 0);
    s.anySlowGroup = false;
    s.nFill = 0;
    auto is_slow = [&](int code) {
        if (code >= 0) return false;
        const int sz = (int)s.groups[-code - 1].mem.size();
        return sz > kFastGroupMax || n > 65535;
    };
    // within a level (its units are independent): the fast units first, and
    // among them the three-member units, then the two-member ones, each with
    // the units that defer their last member's terms first -- so that the
    // lanes of a wave mostly take the same path through the unit code
    auto unit_class = [&](int code) {
        if (is_slow(code)) return 4;
        int sz = 2, dl = -1;
        if (code >= 0) dl = s.defer[2 * (code + 1)];
        else {
            sz = (int)s.groups[-code - 1].mem.size();
            dl = s.groups[-code - 1].deferLast;
        }
        return (sz == 3 ? 0 : 2) + (dl >= 0 ? 0 : 1);
    };
    for (int l = 0; l + 1 < (int)s.levelStart.size(); ++l)
        std::stable_sort(s.units.begin() + s.levelStart[l], s.units.begin() + s.levelStart[l + 1],
                         [&](int a, int b) { return unit_class(a) < unit_class(b); });
    for (size_t k = 0; k < s.units.size(); ++k) {
        int32_t* d = &s.udesc[kDescInts * k];
        const int code = s.units[k];
        double a[3] = {0.0, 0.0, 0.0};
        if (code >= 0) {
            // linear link: a two-member unit (x, x+1) with coefficient HS[4x+1]
            const int x = code;
            if (s.defer[2 * (x + 1)] >= 0 && s.defer[2 * (x + 1) + 1] != 1) {
                err = "internal: linear feeder with more than one term";
                return HSG_E_UNSUPPORTED;
            }
            d[0] = 2;
            d[1] = s.defer[2 * (x + 1)];
            d[2] = -1;
            d[3] = (x & 0xffff) | ((x + 1) << 16);
            d[4] = x + 1;
            d[5] = -1;
            a[0] = s.Off[x];
            if (n > 65535) { err = "cells above 65535 compartments are not supported"; return HSG_E_UNSUPPORTED; }
        } else {
            const int g = -code - 1;
            const auto& G = s.groups[g];
            const int sz = (int)G.mem.size();
            if (is_slow(code)) {
                s.anySlowGroup = true;
                // four or five members: everything the register form needs in
                // the descriptor (members, junction-coefficient offset, the
                // deferred term slots of members 1..s-2)
                bool inDesc = sz <= kDescGroupMax;
                int32_t dinfo[3] = {-1, -1, -1};
                for (int p = 1; inDesc && p + 1 < sz; ++p) {
                    const int ds = s.defer[2 * G.mem[p]], cnt = s.defer[2 * G.mem[p] + 1];
                    if (ds < 0) continue;
                    if (ds >= (1 << 24) || cnt > 127) inDesc = false;
                    else dinfo[p - 1] = ds | (cnt << 24);
                }
                if (inDesc) {
                    d[0] = sz | 32;
                    d[1] = G.deferLast;
                    d[2] = G.hjOff;
                    d[3] = (G.mem[0] & 0xffff) | (G.mem[1] << 16);
                    d[4] = (G.mem[2] & 0xffff) | (G.mem[3] << 16);
                    d[5] = sz == 5 ? G.mem[4] : 0;
                    d[6] = dinfo[0];
                    d[7] = dinfo[1];
                    d[8] = dinfo[2];
                    continue;
                }
                d[0] = sz | 16;
                d[1] = g;
                continue;
            }
            int word = sz;
            if (sz == 3) {
                const int cnt = s.defer[2 * G.mem[1] + 1];
                if (cnt > 15) { err = "internal: too many deferred terms"; return HSG_E_UNSUPPORTED; }
                if (cnt > 0 && s.defer[2 * G.mem[1]] != termBase[g]) { err = "internal: deferred slots out of order"; return HSG_E_UNSUPPORTED; }
                word |= cnt << 8;
            }
            d[0] = word;
            d[1] = G.deferLast;
            d[2] = termBase[g];
            d[3] = (G.mem[0] & 0xffff) | (G.mem[1] << 16);
            d[4] = sz == 3 ? G.mem[2] : G.mem[1];
            d[5] = sz == 3 ? s.nFill++ : -1;
            // the junction coefficients as HJCopy_ holds them: per pair (a,b)
            // of the sorted members, -Ga_a*Ga_b/gsum
            if (sz == 2) a[0] = s.HJ0[G.hjOff];
            else { a[0] = s.HJ0[G.hjOff]; a[1] = s.HJ0[G.hjOff + 2]; a[2] = s.HJ0[G.hjOff + 4]; }
        }
        std::memcpy(&d[6], a, sizeof a);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Message-passing schedule (used by the split-mode solve kernel).  Every node
// is eliminated by exactly one unit (its "consumer": the unit in which it is
// member 0 or 1) and receives its subtree's terms from at most one unit (its
// "feeder": the unit whose last member it is).  Instead of updating the
// node's row in place, the feeder writes its terms into the node's slot and
// the consumer applies them -- in the same order as the in-place form, so the
// arithmetic is unchanged: member 0's row gets its feeder's pairs, member 1's
// row first the term of member 0 and then its feeder's pairs.  On the way
// back the consumer writes the node's VMid into the same slot for the
// feeder.  A slot is live from the feeder's level to the consumer's level;
// slots are shared by nodes whose intervals do not overlap.
// ---------------------------------------------------------------------------
static void schedule_msg(CellSetup& s) {
    s.msgOk = !s.anySlowGroup && s.n <= 65535;
    s.mdesc.clear();
    s.rowPos.assign(s.n, 0);
    s.nSlots = 0;
    s.rootSlot = -1;
    s.rootCnt = 0;
    if (!s.msgOk) return;
    const int n = s.n;
    const int nU = (int)s.units.size();
    const int nL = (int)s.levelStart.size() - 1;
    std::vector<int> level(nU), feeder(n, -1), consumer(n, -1), cslot(n, -1);
    for (int l = 0; l < nL; ++l)
        for (int u = s.levelStart[l]; u < s.levelStart[l + 1]; ++u) level[u] = l;
    auto dsc = [&](int u) { return &s.udesc[(size_t)kDescInts * u]; };
    auto size_of = [&](int u) { return dsc(u)[0] & 15; };
    auto mem = [&](int u, int k) {
        const int32_t* d = dsc(u);
        if (k == 0) return d[3] & 0xffff;
        if (k == 1) return (d[3] >> 16) & 0xffff;
        return d[4];
    };
    auto last_of = [&](int u) { return size_of(u) == 3 ? mem(u, 2) : mem(u, 1); };
    for (int u = 0; u < nU; ++u) {
        feeder[last_of(u)] = u;
        consumer[mem(u, 0)] = u;
        s.rowPos[mem(u, 0)] = 2 * u;
        if (size_of(u) == 3) {
            consumer[mem(u, 1)] = u;
            s.rowPos[mem(u, 1)] = 2 * u + 1;
        }
    }
    const int root = n - 1;
    s.rowPos[root] = 2 * nU;
    // slot intervals [feeder level, consumer level]; the root's consumer is
    // the step after the last level
    std::vector<std::pair<int, int>> iv;   // (start, node)
    for (int m = 0; m < n; ++m)
        if (feeder[m] >= 0) iv.push_back({level[feeder[m]], m});
    std::sort(iv.begin(), iv.end());
    std::vector<int> slotEnd;              // last level each slot is busy
    for (auto& [start, m] : iv) {
        const int end = m == root ? nL : level[consumer[m]];
        int pick = -1;
        for (int k = 0; k < (int)slotEnd.size(); ++k)
            if (slotEnd[k] < start) { pick = k; break; }
        if (pick < 0) { pick = (int)slotEnd.size(); slotEnd.push_back(end); }
        else slotEnd[pick] = end;
        cslot[m] = pick;
    }
    s.nSlots = (int)slotEnd.size();
    auto cnt_of = [&](int m) { return feeder[m] < 0 ? 0 : (size_of(feeder[m]) == 3 ? 2 : 1); };
    s.mdesc.assign((size_t)kMsgDescInts * nU, 0);
    for (int u = 0; u < nU; ++u) {
        const int32_t* d = dsc(u);
        int32_t* o = &s.mdesc[(size_t)kMsgDescInts * u];
        const int sz = size_of(u);
        const int m0 = mem(u, 0), m1 = sz == 3 ? mem(u, 1) : m0;
        o[0] = sz | cnt_of(m0) << 4 | (sz == 3 ? cnt_of(m1) : 0) << 6;
        o[1] = cslot[m0];
        o[2] = sz == 3 ? cslot[m1] : -1;
        o[3] = cslot[last_of(u)];
        o[4] = (m0 & 0xffff) | (m1 << 16);
        o[5] = 0;
        for (int k = 6; k < 12; ++k) o[k] = d[k];
    }
    s.rootSlot = cslot[root];
    s.rootCnt = cnt_of(root);
}

// ---------------------------------------------------------------------------
// Block form of the schedule (large cells, split mode).  The units form a tree
// (a unit's parent eliminates the row it feeds), so any subtree can be
// eliminated on its own as long as it runs before its ancestors: every row
// still receives the same updates in the same order (the deferred-term slots
// keep the sibling order), and the result is the level schedule's, bit for
// bit.  The largest subtrees that fit kBlockMaxLds bytes of LDS are packed
// into blocks of at most that size: one workgroup per block runs them with its rows and descriptors in LDS
// (rows, deferred terms and fills renumbered locally; the rows and term
// pairs the subtrees' roots feed are exported).  The remaining units (the
// top: the ancestors of the blocks, a few dozen units) form one more such
// block, fed only into the root row, run after every block (forward, the
// root, backward) with the pairs the blocks exported staged in; then the
// blocks back-substitute.
// ---------------------------------------------------------------------------
// (the level table holds kBlockForms starts per level, then the end)
static int64_t block_lds(int64_t rows, int64_t pairs, int64_t fills, int64_t levels, int64_t units, int64_t hj) {
    return 16 * rows + 16 * pairs + 8 * fills + 8 * hj + (4 * (kBlockForms * levels + 1) + 15) / 16 * 16 +
           4 * kDescInts * units;
}

namespace {
struct UnitInfo {
    std::vector<std::array<int, 5>> mem;
    std::vector<int> sz, level, readPairs, fill3, lastOf, hj;
};
}  // namespace

// One block from a level-ordered unit list: rows (eliminated, then `fed`),
// term pairs (read inside, then exported by `roots`, or, for the top, every
// pair staged from HBM), fills, local levels and descriptors.  False when a
// local index would not fit.
static bool build_block(const CellSetup& s, const UnitInfo& U, const std::vector<int>& us,
                        const std::vector<int>& roots, const std::vector<int>& fed, bool top,
                        CellSetup::Block& B) {
    auto dsc = [&](int u) { return &s.udesc[(size_t)kDescInts * u]; };
    std::vector<int32_t> elim;
    for (int u : us)
        for (int p = 0; p + 1 < U.sz[u]; ++p) elim.push_back(U.mem[u][p]);
    std::sort(elim.begin(), elim.end());
    std::map<int, int> local;
    for (size_t k = 0; k < elim.size(); ++k) local[elim[k]] = (int)k;
    B.nRows = (int)elim.size();
    B.rows = elim;
    for (int r : fed) {
        local[r] = (int)B.rows.size();
        B.rows.push_back(r);
    }
    B.nExt = (int)fed.size();
    std::map<int, int> pair;
    auto map_range = [&](int base, int cnt) {
        for (int t = 0; t < cnt; ++t)
            if (!pair.count(base + t)) pair[base + t] = (int)pair.size();
    };
    for (int u : us) {
        const int32_t* d = dsc(u);
        if (d[0] & 32) {
            for (int k = 0; k < 3; ++k)
                if (d[6 + k] >= 0) map_range(d[6 + k] & 0xffffff, (d[6 + k] >> 24) & 0x7f);
        } else if (U.readPairs[u] > 0) {
            map_range(d[2], U.readPairs[u]);
        }
    }
    B.nT = (int)pair.size();
    for (int r : roots) {
        const int32_t* d = dsc(r);
        if (d[1] < 0) continue;
        for (int t = 0; t < U.sz[r] - 1; ++t) {
            pair[d[1] + t] = (int)pair.size();
            B.tExt.push_back(d[1] + t);
        }
    }
    // every unit's output pairs are read by its parent inside (roots: exported)
    for (int u : us)
        if (dsc(u)[1] >= 0 && !pair.count(dsc(u)[1])) return false;
    if (top) {
        B.tImp.assign(pair.size(), 0);
        for (auto& kv : pair) B.tImp[kv.second] = kv.first;
    }
    int lo = U.level[us.front()], hi = U.level[us.back()];
    // local levels: the units' own levels, empty ones dropped; within a
    // level the units by form (two or three members, four, five), each
    // form's start in the level table, so that a wavefront runs one form
    std::vector<int> lvOf(hi - lo + 1, -1);
    int nLv = 0;
    for (int u : us)
        if (lvOf[U.level[u] - lo] < 0) lvOf[U.level[u] - lo] = nLv++;
    auto form = [&](int u) { return (dsc(u)[0] & 32) ? U.sz[u] - 3 : 0; };   // 0, or 1-2: register-form groups
    B.level.assign((size_t)kBlockForms * nLv + 1, 0);
    for (int u : us) B.level[(size_t)kBlockForms * lvOf[U.level[u] - lo] + form(u) + 1]++;
    for (size_t k = 0; k + 1 < B.level.size(); ++k) B.level[k + 1] += B.level[k];
    B.desc.assign((size_t)kDescInts * us.size(), 0);
    std::vector<int> pos(B.level.begin(), B.level.end() - 1);
    for (int u : us) {
        int32_t* o = &B.desc[(size_t)kDescInts * pos[(size_t)kBlockForms * lvOf[U.level[u] - lo] + form(u)]++];
        const int32_t* d = dsc(u);
        std::memcpy(o, d, sizeof(int32_t) * kDescInts);
        auto L = [&](int m) { return local.at(m); };
        const auto& m = U.mem[u];
        o[1] = d[1] >= 0 ? pair.at(d[1]) : -1;
        if (d[0] & 32) {
            // the group's junction coefficients, renumbered into the block's
            o[2] = (int)B.hj.size();
            for (int k = 0; k < U.sz[u] * (U.sz[u] - 1); ++k) B.hj.push_back(d[2] + k);
            o[3] = (L(m[0]) & 0xffff) | (L(m[1]) << 16);
            o[4] = (L(m[2]) & 0xffff) | (L(m[3]) << 16);
            o[5] = U.sz[u] == 5 ? L(m[4]) : 0;
            for (int k = 0; k < 3; ++k)
                if (d[6 + k] >= 0) o[6 + k] = pair.at(d[6 + k] & 0xffffff) | (d[6 + k] & ~0xffffff);
        } else {
            o[2] = U.readPairs[u] > 0 ? pair.at(d[2]) : 0;
            o[3] = (L(m[0]) & 0xffff) | (L(m[1]) << 16);
            o[4] = U.sz[u] == 3 ? L(m[2]) : L(m[1]);
            if (U.sz[u] == 3) {
                o[5] = (int)B.fills.size();
                B.fills.push_back(d[5]);
            } else {
                o[5] = -1;
            }
        }
    }
    const int nr = (int)B.rows.size(), np = (int)pair.size(), nf = (int)B.fills.size();
    B.ldsBytes = block_lds(nr, np, nf, nLv, (int64_t)us.size(), (int64_t)B.hj.size());
    // every local index in range (rows, term pairs, fills): the block kernels
    // read and write exactly these
    bool ok = nr <= 65535 && B.ldsBytes <= kBlockMaxLds;
    for (size_t k = 0; ok && k < us.size(); ++k) {
        const int32_t* o = &B.desc[(size_t)kDescInts * k];
        const int S = o[0] & 15;
        int mm[5] = {o[3] & 0xffff, (o[3] >> 16) & 0xffff, 0, 0, 0};
        if (o[0] & 32) {
            mm[2] = o[4] & 0xffff; mm[3] = (o[4] >> 16) & 0xffff; mm[4] = o[5];
            for (int q = 0; q < 3; ++q)
                if (o[6 + q] >= 0) ok = ok && (o[6 + q] & 0xffffff) + ((o[6 + q] >> 24) & 0x7f) <= np;
        } else {
            mm[2] = o[4];
            if (S == 3) ok = ok && o[5] >= 0 && o[5] < nf && o[2] + ((o[0] >> 8) & 15) <= np;
        }
        for (int q = 0; q < S; ++q) ok = ok && mm[q] >= 0 && mm[q] < nr;
        ok = ok && ((o[0] & 32) ? (S == 4 || S == 5) : S <= 3);
        if (o[1] >= 0) ok = ok && o[1] + S - 1 <= np;
    }
    return ok;
}

static void schedule_blocks(CellSetup& s) {
    s.blocks.clear();
    const int n = s.n;
    const int nU = (int)s.units.size();
    if (n < kAutoSplitMin || nU == 0) return;
    const int nL = (int)s.levelStart.size() - 1;
    auto dsc = [&](int u) { return &s.udesc[(size_t)kDescInts * u]; };
    UnitInfo U;
    U.mem.resize(nU);
    U.sz.assign(nU, 0);
    U.level.assign(nU, 0);
    U.readPairs.assign(nU, 0);
    U.fill3.assign(nU, 0);
    U.lastOf.assign(nU, 0);
    U.hj.assign(nU, 0);
    for (int l = 0; l < nL; ++l)
        for (int u = s.levelStart[l]; u < s.levelStart[l + 1]; ++u) U.level[u] = l;
    for (int u = 0; u < nU; ++u) {
        const int32_t* d = dsc(u);
        if (d[0] & 16) return;   // a group kept in memory: no block form
        const int S = d[0] & 15;
        U.sz[u] = S;
        auto& m = U.mem[u];
        m = {d[3] & 0xffff, (d[3] >> 16) & 0xffff, 0, 0, 0};
        if (d[0] & 32) {
            m[2] = d[4] & 0xffff;
            m[3] = (d[4] >> 16) & 0xffff;
            m[4] = d[5];
            for (int k = 0; k < 3; ++k)
                if (d[6 + k] >= 0) U.readPairs[u] += (d[6 + k] >> 24) & 0x7f;
            U.hj[u] = S * (S - 1);
        } else if (S == 3) {
            m[2] = d[4];
            U.readPairs[u] = (d[0] >> 8) & 15;
            U.fill3[u] = 1;
        }
        U.lastOf[u] = m[S - 1];
    }
    std::vector<int> elimUnit(n, -1), parent(nU);
    for (int u = 0; u < nU; ++u)
        for (int p = 0; p + 1 < U.sz[u]; ++p) elimUnit[U.mem[u][p]] = u;
    for (int u = 0; u < nU; ++u) parent[u] = U.lastOf[u] == n - 1 ? -1 : elimUnit[U.lastOf[u]];
    for (int u = 0; u < nU; ++u)
        if (parent[u] >= 0 && U.level[parent[u]] <= U.level[u]) return;   // not a unit tree in level order
    // subtree totals (children come before parents in level order)
    std::vector<int64_t> rowsSub(nU, 0), pairsSub(nU, 0), fillsSub(nU, 0), unitsSub(nU, 0), hjSub(nU, 0);
    for (int u = 0; u < nU; ++u) {
        rowsSub[u] += U.sz[u] - 1;
        pairsSub[u] += U.readPairs[u];
        fillsSub[u] += U.fill3[u];
        unitsSub[u] += 1;
        hjSub[u] += U.hj[u];
        if (parent[u] >= 0) {
            rowsSub[parent[u]] += rowsSub[u];
            pairsSub[parent[u]] += pairsSub[u];
            fillsSub[parent[u]] += fillsSub[u];
            unitsSub[parent[u]] += unitsSub[u];
            hjSub[parent[u]] += hjSub[u];
        }
    }
    // a subtree fits a block when its rows, the row it feeds, its term pairs
    // (read inside, then exported), fills, levels, descriptors and junction
    // coefficients fit the block's LDS
    auto fits = [&](int u) {
        return block_lds(rowsSub[u] + 1, pairsSub[u] + (U.sz[u] - 1), fillsSub[u], U.level[u] + 1, unitsSub[u],
                         hjSub[u]) <= kBlockMaxLds;
    };
    std::vector<int> roots;
    for (int u = 0; u < nU; ++u)
        if (fits(u) && (parent[u] < 0 || !fits(parent[u]))) {
            if (parent[u] < 0) return;   // the whole cell is one block: nothing to gain
            roots.push_back(u);
        }
    if (roots.size() < 2) return;
    std::vector<int> sub(nU, -1);
    for (int u : roots) sub[u] = u;
    for (int u = nU - 1; u >= 0; --u)
        if (sub[u] < 0 && parent[u] >= 0 && sub[parent[u]] >= 0) sub[u] = sub[parent[u]];
    // pack subtrees into blocks, in the order of the rows they feed
    std::sort(roots.begin(), roots.end(), [&](int a, int b) { return U.lastOf[a] < U.lastOf[b]; });
    std::vector<int> packOf(nU, -1);
    std::vector<std::vector<int>> packs;
    {
        int64_t r = 0, pr = 0, f = 0, lv = 0, nu = 0, hj = 0;
        for (int u : roots) {
            const int64_t nr = r + rowsSub[u] + 1, np = pr + pairsSub[u] + (U.sz[u] - 1), nf = f + fillsSub[u];
            const int64_t nlv = std::max<int64_t>(lv, U.level[u] + 1), nnu = nu + unitsSub[u], nhj = hj + hjSub[u];
            if (packs.empty() || block_lds(nr, np, nf, nlv, nnu, nhj) > kBlockMaxLds) {
                packs.push_back({});
                r = rowsSub[u] + 1;
                pr = pairsSub[u] + (U.sz[u] - 1);
                f = fillsSub[u];
                lv = U.level[u] + 1;
                nu = unitsSub[u];
                hj = hjSub[u];
            } else {
                r = nr; pr = np; f = nf; lv = nlv; nu = nnu; hj = nhj;
            }
            packs.back().push_back(u);
            packOf[u] = (int)packs.size() - 1;
        }
    }
    std::vector<std::vector<int>> unitsOf(packs.size());
    std::vector<int> topUnits;
    for (int u = 0; u < nU; ++u) {
        if (sub[u] >= 0) unitsOf[packOf[sub[u]]].push_back(u);   // level order kept
        else topUnits.push_back(u);
    }
    s.blocks.resize(packs.size() + 1);
    for (size_t b = 0; b < packs.size(); ++b) {
        std::vector<int> fed;
        for (int r : packs[b]) fed.push_back(U.lastOf[r]);
        if (!build_block(s, U, unitsOf[b], packs[b], fed, false, s.blocks[b])) { s.blocks.clear(); return; }
    }
    if (topUnits.empty() ||
        !build_block(s, U, topUnits, {}, {n - 1}, true, s.blocks.back())) {
        s.blocks.clear();
        return;
    }
}

int setup_cell(const hsg_cell_desc* d, double dt, int caAdvance, CellSetup& s, std::string& err) {
    if (!d) { err = "null cell description"; return HSG_E_ARG; }
    if (!(dt > 0.0)) { err = "HSolve: 'dt' must be set (positive) before 'target'"; return HSG_E_ARG; }
    if (caAdvance != 0 && caAdvance != 1) { err = "HSolve: caAdvance should be either 0 or 1"; return HSG_E_ARG; }
    const int nd = d->n_compartments;
    if (nd <= 0 || !d->compartments) { err = "no compartments"; return HSG_E_MODEL; }
    if (d->seed < 0 || d->seed >= nd) { err = "seed compartment out of range"; return HSG_E_ARG; }
    for (int k = 0; k < d->n_axial; ++k) {
        int p = d->axial[2 * k], c = d->axial[2 * k + 1];
        if (p < 0 || p >= nd || c < 0 || c >= nd) { err = "axial message refers to a missing compartment"; return HSG_E_ARG; }
    }
    s.dt = dt;
    s.caAdvance = caAdvance;
    Adjacency adj = make_adjacency(d);
    int rc = walk_tree(d, adj, s.hinesToDesc, err);
    if (rc) return rc;
    const int n = (int)s.hinesToDesc.size();
    s.n = n;
    s.descToHines.assign(nd, -1);
    for (int i = 0; i < n; ++i) s.descToHines[s.hinesToDesc[i]] = i;

    // initialize + storeTree
    std::vector<std::vector<int>> leaks(nd);
    for (int k = 0; k < d->n_leaks; ++k) {
        int c = d->leaks[k].compartment;
        if (c < 0 || c >= nd) { err = "leak on a missing compartment"; return HSG_E_ARG; }
        leaks[c].push_back(k);
    }
    s.V.resize(n); s.CmByDt.resize(n); s.EmByRm.resize(n);
    s.Ra.resize(n); s.Rm.resize(n); s.Cm.resize(n); s.Em.resize(n); s.initVm.resize(n);
    s.injBasal.assign(n, 0.0); s.hasInject.assign(n, 0);
    std::vector<std::vector<int>> children(n);
    for (int i = 0; i < n; ++i) {
        const int dc = s.hinesToDesc[i];
        const hsg_compartment_desc& c = d->compartments[dc];
        s.V[i] = c.Vm;
        double EmGmThev = c.Em / c.Rm;
        for (int k : leaks[dc]) EmGmThev += d->leaks[k].Ek * d->leaks[k].Gk;
        s.CmByDt[i] = 2.0 * c.Cm / dt;
        s.EmByRm[i] = EmGmThev;
        if (c.inject != 0.0) { s.hasInject[i] = 1; s.injBasal[i] = c.inject; }
        for (int ch : adj.axial[dc]) {
            int h = s.descToHines[ch];
            if (h < 0) { err = "child compartment outside the walked tree"; return HSG_E_MODEL; }
            children[i].push_back(h);
        }
        s.Ra[i] = c.Ra; s.Rm[i] = c.Rm; s.Cm[i] = c.Cm; s.Em[i] = c.Em; s.initVm[i] = c.initVm;
    }

    // HinesMatrix::setup: Ga, makeJunctions, makeMatrix
    std::vector<double> Ga(n);
    for (int i = 0; i < n; ++i) Ga[i] = 2.0 / s.Ra[i];
    std::vector<std::vector<int>> coupled;
    for (int i = 0; i < n; ++i) {
        const auto& c = children[i];
        if (c.empty()) continue;
        if (c.size() == 1 && (c[0] - i == 1 || c[0] - i == -1)) continue;
        std::vector<int> g = c;
        g.push_back(i);
        std::sort(g.begin(), g.end());
        coupled.push_back(std::move(g));
    }
    std::stable_sort(coupled.begin(), coupled.end(),
                     [](const std::vector<int>& a, const std::vector<int>& b) { return a[0] < b[0]; });
    std::vector<int> junctionGroup(n, -1);
    for (int g = 0; g < (int)coupled.size(); ++g)
        for (size_t p = 0; p + 1 < coupled[g].size(); ++p) junctionGroup[coupled[g][p]] = g;

    std::vector<double> HS2(n);
    s.Off.assign(n, 0.0);
    for (int i = 0; i < n; ++i) HS2[i] = s.Cm[i] / (dt / 2.0) + 1.0 / s.Rm[i];
    for (int i = 0; i + 1 < n; ++i) {
        if (junctionGroup[i] >= 0) continue;
        double gi = Ga[i], gj = Ga[i + 1];
        double gij = gi * gj / (gi + gj);
        s.Off[i] = -gij;
        HS2[i] += gij;
        HS2[i + 1] += gij;
    }
    for (const auto& g : coupled) {
        double gsum = 0.0;
        for (int x : g) gsum += Ga[x];
        for (int x : g) { double gi = Ga[x]; HS2[x] += gi * (1.0 - gi / gsum); }
    }
    s.Dbase = HS2;
    s.HJ0.clear();
    s.groups.clear();
    s.maxGroup = 0;
    for (const auto& g : coupled) {
        double gsum = 0.0;
        for (int x : g) gsum += Ga[x];
        Group G;
        G.mem = g;
        G.hjOff = (int)s.HJ0.size();
        G.deferLast = -1;
        for (size_t a = 0; a + 1 < g.size(); ++a)
            for (size_t b = a + 1; b < g.size(); ++b) {
                double gij = Ga[g[a]] * Ga[g[b]] / gsum;
                s.HJ0.push_back(-gij);
                s.HJ0.push_back(-gij);
            }
        s.maxGroup = std::max(s.maxGroup, (int)g.size());
        s.groups.push_back(std::move(G));
    }
    rc = schedule(s, junctionGroup, err);
    if (rc) return rc;
    schedule_msg(s);
    schedule_blocks(s);

    // readHHChannels / readGates
    std::vector<std::vector<int>> chans(nd);
    for (int k = 0; k < d->n_channels; ++k) {
        int c = d->channels[k].compartment;
        if (c < 0 || c >= nd) { err = "channel on a missing compartment"; return HSG_E_ARG; }
        chans[c].push_back(k);
    }
    const int nch = d->n_channels;
    s.chanCompt.assign(nch, -1); s.chanSlot.assign(nch, -1); s.chanState.assign(nch, -1);
    s.chanXp.assign(nch, 0); s.chanYp.assign(nch, 0); s.chanZp.assign(nch, 0);
    int maxSlots = 0, maxGates = 0;
    std::vector<int> chanOrder;                 // channelId_ order
    for (int i = 0; i < n; ++i) {
        const auto& cl = chans[s.hinesToDesc[i]];
        maxSlots = std::max(maxSlots, (int)cl.size());
        int g = 0;
        for (int k : cl) {
            const hsg_channel_desc& c = d->channels[k];
            g += (c.Xpower > 0.0) + (c.Ypower > 0.0) + (c.Zpower > 0.0);
            chanOrder.push_back(k);
        }
        maxGates = std::max(maxGates, g);
    }
    if (maxSlots > 8) { err = "more than 8 HHChannels on one compartment"; return HSG_E_UNSUPPORTED; }
    if (maxGates > 24) { err = "more than 24 gates on one compartment"; return HSG_E_UNSUPPORTED; }
    s.chanSlots = maxSlots;
    s.gateSlots = maxGates;
    // gates in channelId_ order, gCaDepend_ marks a channel's last gate
    std::vector<int> gateOf;        // desc gate index per gate entry
    std::vector<char> caDep;
    std::vector<int> gateChan, gateWhich;
    for (int k : chanOrder) {
        const hsg_channel_desc& c = d->channels[k];
        const double pw[3] = {c.Xpower, c.Ypower, c.Zpower};
        for (int w = 0; w < 3; ++w)
            if (pw[w] > 0.0) {
                if (c.gate[w] < 0 || c.gate[w] >= d->n_gates || !d->gates) {
                    err = "channel gate with power > 0 has no HHGate";
                    return HSG_E_MODEL;
                }
                gateOf.push_back(c.gate[w]);
                caDep.push_back(0);
                gateChan.push_back(k);
                gateWhich.push_back(w);
            }
        if (c.useConcentration && !caDep.empty()) caDep.back() = 1;
    }

    // readCalcium
    const int npd = d->n_pools;
    s.poolCompt.assign(npd, -1);
    s.poolSlot.assign(npd, -1);
    s.poolCount.assign(n, 0);
    std::vector<int> caTgtLocal(nch, -1), caDepLocal(nch, -1);
    int maxPools = 0;
    {
        std::vector<int> caConcIndex(npd, -1);
        for (int i = 0; i < n; ++i) {
            for (int k : chans[s.hinesToDesc[i]]) {
                const hsg_channel_desc& c = d->channels[k];
                int ids[2], nid = 0;
                if (c.caTarget >= 0) ids[nid++] = c.caTarget;
                if (c.caDepend >= 0) ids[nid++] = c.caDepend;
                for (int q = 0; q < nid; ++q) {
                    int p = ids[q];
                    if (p >= npd) { err = "channel refers to a missing CaConc"; return HSG_E_ARG; }
                    if (caConcIndex[p] < 0) {
                        caConcIndex[p] = s.poolCount[i]++;
                        s.poolCompt[p] = i;
                        s.poolSlot[p] = caConcIndex[p];
                    } else if (s.poolCompt[p] != i) {
                        err = "a CaConc pool shared by channels of two compartments";
                        return HSG_E_UNSUPPORTED;
                    }
                }
                if (c.caTarget >= 0) caTgtLocal[k] = caConcIndex[c.caTarget];
                if (c.caDepend >= 0) caDepLocal[k] = caConcIndex[c.caDepend];
            }
            maxPools = std::max(maxPools, s.poolCount[i]);
        }
    }
    if (maxPools > 4) { err = "more than 4 CaConc pools on one compartment"; return HSG_E_UNSUPPORTED; }
    s.poolSlots = maxPools;
    const size_t PS = (size_t)maxPools * n;
    s.pC.assign(PS, 0.0); s.pCa.assign(PS, 0.0); s.pBasal.assign(PS, 0.0);
    s.pF1.assign(PS, 0.0); s.pF2.assign(PS, 0.0); s.pCeil.assign(PS, 0.0); s.pFloor.assign(PS, 0.0);
    for (int p = 0; p < npd; ++p) {
        if (s.poolCompt[p] < 0) continue;
        const hsg_caconc_desc& pc = d->pools[p];
        const size_t o = (size_t)s.poolSlot[p] * n + s.poolCompt[p];
        // CaConcStruct(Ca, CaBasal, tau, B, ceiling, floor, dt) (HSolveStruct.cpp:103-136)
        s.pBasal[o] = pc.CaBasal;
        s.pC[o] = pc.Ca - pc.CaBasal;
        s.pF1[o] = 4.0 / (2.0 + dt / pc.tau) - 1.0;
        s.pF2[o] = 2.0 * pc.B * dt / (2.0 + dt / pc.tau);
        s.pCeil[o] = pc.ceiling;
        s.pFloor[o] = pc.floor;
        s.pCa[o] = 0.0;   // ca_ stays zero until reinit
    }

    // createLookupTables
    std::vector<int> vGate, caGate;
    for (size_t g = 0; g < gateOf.size(); ++g) (caDep[g] ? caGate : vGate).push_back(gateOf[g]);
    auto uniq = [&](std::vector<int>& v) {
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return d->gates[a].id < d->gates[b].id; });
        std::vector<int> u;
        for (int x : v) if (u.empty() || d->gates[u.back()].id != d->gates[x].id) u.push_back(x);
        v = u;
    };
    uniq(vGate);
    uniq(caGate);
    std::map<int64_t, int> species;
    for (size_t i = 0; i < caGate.size(); ++i) species[d->gates[caGate[i]].id] = (int)i;
    for (size_t i = 0; i < vGate.size(); ++i) species[d->gates[vGate[i]].id] = (int)i;
    auto grid = [&](const std::vector<int>& gs, double& mn, double& mx, int& div) {
        mn = DBL_MAX;
        mx = DBL_MIN;
        double ddx = DBL_MAX;
        for (int g : gs) {
            const hsg_gate_desc& G = d->gates[g];
            if (G.divs == 0 || !G.tableA || !G.tableB) return false;
            double dx = (G.max - G.min) / G.divs;
            if (G.min < mn) mn = G.min;
            if (G.max > mx) mx = G.max;
            if (dx < ddx) ddx = dx;
        }
        double dv = (mx - mn) / ddx;
        div = (int)(dv + 0.5);
        return true;
    };
    double vMin, vMax, caMin, caMax;
    int vDiv, caDiv;
    if (!grid(caGate, caMin, caMax, caDiv) || !grid(vGate, vMin, vMax, vDiv)) {
        err = "HHGate without tables";
        return HSG_E_MODEL;
    }
    if (vDiv < 0 || vDiv > (1 << 22) || caDiv < 0 || caDiv > (1 << 22)) {
        err = "lookup table grid out of range";
        return HSG_E_MODEL;
    }
    make_table(s.tables.ca, caMin, caMax, (unsigned)caDiv, (unsigned)caGate.size());
    make_table(s.tables.v, vMin, vMax, (unsigned)vDiv, (unsigned)vGate.size());
    std::vector<double> A, B;
    for (size_t i = 0; i < caGate.size(); ++i) {
        rates(d->gates[caGate[i]], caMin, caMax, (unsigned)caDiv, A, B);
        add_columns(s.tables.ca, (int)i, A, B);
    }
    for (size_t i = 0; i < vGate.size(); ++i) {
        rates(d->gates[vGate[i]], vMin, vMax, (unsigned)vDiv, A, B);
        add_columns(s.tables.v, (int)i, A, B);
    }
    {
        uint64_t h = 1469598103934665603ull;
        for (const HostTable* t : {&s.tables.v, &s.tables.ca}) {
            h = fnv(&t->min, sizeof(double), h);
            h = fnv(&t->max, sizeof(double), h);
            h = fnv(&t->nDivs, sizeof(uint32_t), h);
            h = fnv(&t->nSpecies, sizeof(uint32_t), h);
            h = fnv(t->table.data(), t->table.size() * sizeof(double), h);
        }
        char buf[64];
        snprintf(buf, sizeof buf, "%016llx:%zu", (unsigned long long)h, s.tables.v.table.size());
        s.tables.key = buf;
    }
    // column of each gate entry, then per channel X/Y/Z columns
    std::vector<int> colX(nch, 0), colY(nch, 0), colZ(nch, 0);
    for (size_t g = 0; g < gateOf.size(); ++g) {
        int col = 2 * species[d->gates[gateOf[g]].id];
        int k = gateChan[g];
        if (gateWhich[g] == 0) colX[k] = col;
        else if (gateWhich[g] == 1) colY[k] = col;
        else colZ[k] = col;
    }
    // the reference reads the vTable at a Ca column when a Ca-dependent Z gate
    // has no pool and no external Ca (HSolveActive.cpp:304-312); that read
    // must stay inside the vTable
    for (int k = 0; k < nch; ++k) {
        const hsg_channel_desc& c = d->channels[k];
        if (c.Zpower > 0.0 && caDepLocal[k] < 0 && colZ[k] + 1 >= (int)s.tables.v.nColumns() &&
            !(colZ[k] + 1 < (int)s.tables.ca.nColumns())) {
            err = "Ca-dependent Z gate with no pool: its column lies outside both tables";
            return HSG_E_UNSUPPORTED;
        }
        if (c.Zpower > 0.0 && caDepLocal[k] >= 0 && colZ[k] + 1 >= (int)s.tables.ca.nColumns()) {
            err = "Z gate reads a CaConc but its column lies outside the Ca table (useConcentration unset?)";
            return HSG_E_UNSUPPORTED;
        }
        if (c.Xpower > 0.0 && colX[k] + 1 >= (int)s.tables.v.nColumns()) {
            err = "X gate column outside the voltage table (Ca-dependent X gate)";
            return HSG_E_UNSUPPORTED;
        }
        if (c.Ypower > 0.0 && colY[k] + 1 >= (int)s.tables.v.nColumns()) {
            err = "Y gate column outside the voltage table (Ca-dependent Y gate)";
            return HSG_E_UNSUPPORTED;
        }
    }

    // channel slots, local types
    const size_t CS = (size_t)maxSlots * n;
    s.chTypeLocal.assign(CS, -1);
    s.Gbar.assign(CS, 0.0); s.Ek.assign(CS, 0.0); s.mod.assign(CS, 1.0);
    s.state.assign((size_t)maxGates * n, 0.0);
    s.localTypes.clear();
    std::map<std::string, int> typeIds;
    s.hasModulation = false;
    for (int i = 0; i < n; ++i) {
        int k = 0, js = 0;
        for (int dk : chans[s.hinesToDesc[i]]) {
            const hsg_channel_desc& c = d->channels[dk];
            HostChanType t;
            t.Xpower = c.Xpower; t.Ypower = c.Ypower; t.Zpower = c.Zpower;
            t.instant = c.instant;
            t.nGates = (c.Xpower > 0.0) + (c.Ypower > 0.0) + (c.Zpower > 0.0);
            t.colX = c.Xpower > 0.0 ? colX[dk] : 0;
            t.colY = c.Ypower > 0.0 ? colY[dk] : 0;
            t.colZ = c.Zpower > 0.0 ? colZ[dk] : 0;
            t.caDepLocal = c.Zpower > 0.0 ? caDepLocal[dk] : -1;
            t.caTgtLocal = caTgtLocal[dk];
            t.tableKey = s.tables.key;
            char kb[256];
            snprintf(kb, sizeof kb, "%a|%a|%a|%d|%d|%d|%d|%d|%d", t.Xpower, t.Ypower, t.Zpower,
                     t.instant, t.colX, t.colY, t.colZ, t.caDepLocal, t.caTgtLocal);
            auto it = typeIds.find(kb);
            int tid;
            if (it == typeIds.end()) {
                tid = (int)s.localTypes.size();
                typeIds[kb] = tid;
                s.localTypes.push_back(t);
            } else tid = it->second;
            const size_t o = (size_t)k * n + i;
            s.chTypeLocal[o] = tid;
            s.Gbar[o] = c.Gbar;
            s.Ek[o] = c.Ek;
            s.mod[o] = c.modulation;
            if (c.modulation != 1.0) s.hasModulation = true;
            s.chanCompt[dk] = i;
            s.chanSlot[dk] = k;
            s.chanState[dk] = js;
            s.chanXp[dk] = c.Xpower; s.chanYp[dk] = c.Ypower; s.chanZp[dk] = c.Zpower;
            if (c.Xpower > 0.0) s.state[(size_t)(js++) * n + i] = c.X;
            if (c.Ypower > 0.0) s.state[(size_t)(js++) * n + i] = c.Y;
            if (c.Zpower > 0.0) s.state[(size_t)(js++) * n + i] = c.Z;
            ++k;
        }
    }
    // uniform cell: every compartment carries the same type in every slot
    s.uniform = true;
    s.slotTypeUniform.assign(maxSlots, -1);
    for (int k = 0; k < maxSlots && s.uniform; ++k) {
        int t0 = s.chTypeLocal[(size_t)k * n];
        for (int i = 1; i < n; ++i)
            if (s.chTypeLocal[(size_t)k * n + i] != t0) { s.uniform = false; break; }
        s.slotTypeUniform[k] = t0;
    }

    // HHChannel2D objects (biophysics/HHChannel2D.cpp) and their tables
    s.tables2d.clear();
    for (int k = 0; k < (d->n_tables2d > 0 ? d->n_tables2d : 0); ++k) {
        const hsg_table2d_desc& td = d->tables2d[k];
        if (!td.table) { err = "2-D table without values"; return HSG_E_ARG; }
        HostTable2D t;
        t.nx = td.xdivs + 1;
        t.ny = td.ydivs + 1;
        t.xmin = td.xmin; t.xmax = td.xmax; t.ymin = td.ymin; t.ymax = td.ymax;
        t.invDx = td.xmax != td.xmin ? td.xdivs / (td.xmax - td.xmin) : 1.0;
        t.invDy = td.ymax != td.ymin ? td.ydivs / (td.ymax - td.ymin) : 1.0;
        t.table.assign(td.table, td.table + (size_t)t.nx * t.ny);
        uint64_t h = 1469598103934665603ULL;
        h = fnv(&t.nx, sizeof t.nx, h);
        h = fnv(&t.ny, sizeof t.ny, h);
        h = fnv(&t.xmin, 4 * sizeof(double), h);
        h = fnv(t.table.data(), t.table.size() * sizeof(double), h);
        t.hash = h;
        s.tables2d.push_back(std::move(t));
    }
    const int n2 = d->n_chans2d > 0 ? d->n_chans2d : 0;
    std::vector<std::vector<int>> byCompt(n);
    for (int k = 0; k < n2; ++k) {
        const hsg_chan2d_desc& cd = d->chans2d[k];
        if (cd.compartment < 0 || cd.compartment >= nd || s.descToHines[cd.compartment] < 0) {
            err = "HHChannel2D on a compartment outside the cell";
            return HSG_E_MODEL;
        }
        byCompt[s.descToHines[cd.compartment]].push_back(k);
    }
    s.ch2d.clear();
    s.ch2dStart.assign(n + 1, 0);
    s.ch2dOfDesc.assign(n2, -1);
    for (int i = 0; i < n; ++i) {
        s.ch2dStart[i] = (int32_t)s.ch2d.size();
        for (int k : byCompt[i]) {
            const hsg_chan2d_desc& cd = d->chans2d[k];
            HostChan2D c{};
            c.compt = i;
            c.desc = k;
            c.instant = cd.instant;
            const double pw[3] = {cd.Xpower, cd.Ypower, cd.Zpower};
            const double st[3] = {cd.X, cd.Y, cd.Z};
            for (int a = 0; a < 3; ++a) {
                c.pw[a] = pw[a];
                c.st[a] = st[a];
                c.tabA[a] = cd.tableA[a];
                c.tabB[a] = cd.tableB[a];
                if (pw[a] > 0.0 && (cd.tableA[a] < 0 || cd.tableA[a] >= (int)s.tables2d.size() ||
                                    cd.tableB[a] < 0 || cd.tableB[a] >= (int)s.tables2d.size())) {
                    err = "HHChannel2D gate without its tables";
                    return HSG_E_MODEL;
                }
            }
            for (int q = 0; q < 6; ++q) {
                if (cd.dep[q] < -1 || cd.dep[q] > 2) { err = "HHChannel2D: unknown table dependency"; return HSG_E_ARG; }
                c.dep[q] = cd.dep[q];
            }
            auto pool_at = [&](int p, int& out) {
                out = -1;
                if (p < 0) return true;
                if (p >= npd || s.poolCompt[p] < 0) return false;   // a pool the solver does not run
                out = s.poolSlot[p] * n + s.poolCompt[p];
                return true;
            };
            if (!pool_at(cd.conc1, c.conc1) || !pool_at(cd.conc2, c.conc2)) {
                err = "HHChannel2D reads a CaConc that no HHChannel of the cell uses";
                return HSG_E_UNSUPPORTED;
            }
            c.Gbar = cd.Gbar;
            c.Ek = cd.Ek;
            c.mod = cd.modulation;
            s.ch2dOfDesc[k] = (int)s.ch2d.size();
            s.ch2d.push_back(c);
        }
    }
    s.ch2dStart[n] = (int32_t)s.ch2d.size();

    // PulseGens on the compartments' injectMsg: their targets get an
    // injection entry, as HSolve::addInject creates one
    // (hsolve/HSolveInterface.cpp:213-219)
    s.pulses.clear();
    for (int k = 0; k < (d->n_pulsegens > 0 ? d->n_pulsegens : 0); ++k) {
        const hsg_pulsegen_desc& pd = d->pulsegens[k];
        if (pd.compartment < -1 || pd.compartment >= nd ||
            (pd.compartment >= 0 && s.descToHines[pd.compartment] < 0)) {
            err = "PulseGen on a compartment outside the cell";
            return HSG_E_MODEL;
        }
        if (pd.n_pulses < 1 || pd.n_pulses > HSG_MAX_PULSES) {
            err = "PulseGen: 1 to HSG_MAX_PULSES pulses";
            return HSG_E_ARG;
        }
        if (pd.trigMode < 0 || pd.trigMode > 2) {
            err = "PulseGen: trigMode 0, 1 or 2";
            return HSG_E_ARG;
        }
        const int np = d->n_pulsegens;
        if (pd.trigMode != 0 && (pd.input_src < -1 || pd.input_src >= np || pd.input_src == k)) {
            err = "PulseGen: input_src is -1 or another PulseGen of the cell";
            return HSG_E_ARG;
        }
        HostPulse p{};
        p.compt = pd.compartment >= 0 ? s.descToHines[pd.compartment] : -1;
        p.n = pd.n_pulses;
        p.trigMode = pd.trigMode;
        p.src = pd.trigMode != 0 ? pd.input_src : -1;
        p.base = pd.baseLevel;
        for (int j = 0; j < p.n; ++j) {
            p.delay[j] = pd.delay[j];
            p.width[j] = pd.width[j];
            p.level[j] = pd.level[j];
        }
        if (p.compt >= 0) s.hasInject[p.compt] = 1;
        s.pulses.push_back(p);
    }
    for (const HostPulse& p : s.pulses) {
        if (p.compt < 0 && p.trigMode != 0) {
            err = "PulseGen: one without a compartment can only be a free-running input source";
            return HSG_E_UNSUPPORTED;
        }
        if (p.src >= 0 && !pulse_source_ok(s.pulses[p.src])) {
            err = "PulseGen: an input source must run free with pulses that cover its period";
            return HSG_E_UNSUPPORTED;
        }
    }
    // VClamps likewise (their currentOut goes to injectMsg)
    s.vclamps.clear();
    for (int k = 0; k < (d->n_vclamps > 0 ? d->n_vclamps : 0); ++k) {
        const hsg_vclamp_desc& vd = d->vclamps[k];
        if (vd.compartment < 0 || vd.compartment >= nd || s.descToHines[vd.compartment] < 0) {
            err = "VClamp on a compartment outside the cell";
            return HSG_E_MODEL;
        }
        if (vd.mode < 0 || vd.mode > 2) {
            err = "VClamp: mode must be 0, 1 or 2";
            return HSG_E_ARG;
        }
        HostVClamp v{};
        v.compt = s.descToHines[vd.compartment];
        v.mode = vd.mode;
        v.tau = vd.tau;
        v.ti = vd.ti;
        v.td = vd.td;
        v.gain = vd.gain;
        s.hasInject[v.compt] = 1;
        s.vclamps.push_back(v);
    }
    {
        std::vector<int> entryOf(n, -1);
        int m = 0;
        for (int i = 0; i < n; ++i)
            if (s.hasInject[i]) entryOf[i] = m++;
        for (auto& p : s.pulses) p.entry = p.compt >= 0 ? entryOf[p.compt] : -1;
        for (auto& v : s.vclamps) v.entry = entryOf[v.compt];
        s.nSetupEntries = m;
        // objects by entry (those with no entry are left out)
        auto ranges = [m](const std::vector<int>& entryOfObj, std::vector<int32_t>& start, std::vector<int32_t>& idx) {
            start.assign(m + 1, 0);
            for (int e : entryOfObj)
                if (e >= 0) start[e + 1]++;
            for (int e = 0; e < m; ++e) start[e + 1] += start[e];
            idx.assign(start[m], 0);
            std::vector<int32_t> fill(start.begin(), start.end() - 1);
            for (size_t k = 0; k < entryOfObj.size(); ++k)
                if (entryOfObj[k] >= 0) idx[fill[entryOfObj[k]]++] = (int32_t)k;
        };
        std::vector<int> pe, ve;
        for (auto& p : s.pulses) pe.push_back(p.entry);
        for (auto& v : s.vclamps) ve.push_back(v.entry);
        ranges(pe, s.pulseStart, s.pulseIdx);
        ranges(ve, s.vcStart, s.vcIdx);
    }

    // Leakage objects (their per-tick sends) and SynChans, grouped by
    // compartment in Hines order, desc order within one
    {
        std::vector<std::vector<int>> byC(n);
        for (int k = 0; k < d->n_leaks; ++k) {
            const int h = s.descToHines[d->leaks[k].compartment];
            if (h >= 0) byC[h].push_back(k);   // a leak off the tree is not the solver's
        }
        s.leaks.clear();
        s.leakStart.assign(n + 1, 0);
        s.leakOfDesc.assign(d->n_leaks > 0 ? d->n_leaks : 0, -1);
        for (int i = 0; i < n; ++i) {
            s.leakStart[i] = (int32_t)s.leaks.size();
            for (int k : byC[i]) {
                const hsg_leak_desc& ld = d->leaks[k];
                s.leakOfDesc[k] = (int)s.leaks.size();
                s.leaks.push_back(HostLeak{i, k, ld.Gbar, ld.Ek, ld.modulation});
            }
        }
        s.leakStart[n] = (int32_t)s.leaks.size();
    }
    {
        const int ns = d->n_synchans > 0 ? d->n_synchans : 0;
        if (ns > 0 && !d->synchans) { err = "synchans: null array"; return HSG_E_ARG; }
        std::vector<std::vector<int>> byC(n);
        std::vector<int> first(ns, 0);
        int nsyn = 0;
        for (int k = 0; k < ns; ++k) {
            const hsg_synchan_desc& sd = d->synchans[k];
            if (sd.compartment < 0 || sd.compartment >= nd || s.descToHines[sd.compartment] < 0) {
                err = "SynChan on a compartment outside the cell";
                return HSG_E_MODEL;
            }
            if (sd.n_synapses < 0) { err = "SynChan: negative n_synapses"; return HSG_E_ARG; }
            if (!(sd.tau1 > 0.0) || !(sd.tau2 >= 0.0)) { err = "SynChan: tau1 must be > 0, tau2 >= 0"; return HSG_E_ARG; }
            first[k] = nsyn;
            nsyn += sd.n_synapses;
            byC[s.descToHines[sd.compartment]].push_back(k);
        }
        if (nsyn > 0 && !d->synapses) { err = "synapses: null array"; return HSG_E_ARG; }
        s.synW.resize(nsyn);
        s.synDelay.resize(nsyn);
        for (int j = 0; j < nsyn; ++j) {
            s.synW[j] = d->synapses[j].weight;
            s.synDelay[j] = d->synapses[j].delay;
            if (!std::isfinite(s.synW[j]) || !(s.synDelay[j] >= 0.0) || !std::isfinite(s.synDelay[j])) {
                err = "synapse: weight must be finite and delay finite and >= 0";
                return HSG_E_ARG;
            }
        }
        s.syn.clear();
        s.synStart.assign(n + 1, 0);
        s.synOfDesc.assign(ns, -1);
        for (int i = 0; i < n; ++i) {
            s.synStart[i] = (int32_t)s.syn.size();
            for (int k : byC[i]) {
                const hsg_synchan_desc& sd = d->synchans[k];
                s.synOfDesc[k] = (int)s.syn.size();
                s.syn.push_back(HostSyn{i, k, first[k], sd.n_synapses, sd.Gbar, sd.Ek, sd.tau1, sd.tau2,
                                        sd.modulation});
            }
        }
        s.synStart[n] = (int32_t)s.syn.size();
    }

    // SpikeGens in Hines order of their compartments
    std::vector<std::vector<int>> spk(nd);
    for (int k = 0; k < d->n_spikegens; ++k) {
        int c = d->spikegens[k].compartment;
        if (c < 0 || c >= nd) { err = "SpikeGen on a missing compartment"; return HSG_E_ARG; }
        spk[c].push_back(k);
    }
    for (int i = 0; i < n; ++i)
        for (int k : spk[s.hinesToDesc[i]]) {
            s.spkCompt.push_back(i);
            s.spkDesc.push_back(k);
            s.spkEdge.push_back(d->spikegens[k].edgeTriggered);
            s.spkThr.push_back(d->spikegens[k].threshold);
            s.spkRefr.push_back(d->spikegens[k].refractT);
        }
    return 0;
}

double pulse_free_output(const HostPulse& p, double t) {
    double period = p.width[0] + p.delay[0];
    for (int ii = 1; ii < p.n && (p.width[ii] > 0.0 || p.delay[ii] > 0.0); ++ii) {
        const double incr = p.delay[ii] + p.width[ii] - p.width[ii - 1];
        if (incr > 0) period += incr;
    }
    double phase = std::fmod(t, period);
    for (int ii = 0; ii < p.n; ++ii) {
        if (phase < p.delay[ii]) return p.base;
        if (phase < p.delay[ii] + p.width[ii]) return p.level[ii];
        phase -= p.delay[ii];
    }
    return p.base;
}

bool pulse_source_ok(const HostPulse& p) {
    if (p.trigMode != 0 || p.n < 1) return false;
    // the period as PulseGen::process computes it (device/PulseGen.cpp:362-369)
    double period = p.width[0] + p.delay[0];
    for (int ii = 1; ii < p.n && (p.width[ii] > 0.0 || p.delay[ii] > 0.0); ++ii) {
        const double incr = p.delay[ii] + p.width[ii] - p.width[ii - 1];
        if (incr > 0) period += incr;
    }
    if (!(period > 0.0) || !std::isfinite(period)) return false;
    // a phase escapes every pulse and gap only at or past
    // max_ii (delay[0] + ... + delay[ii] + max(0, width[ii]))
    double D = 0.0, reach = -INFINITY;
    for (int ii = 0; ii < p.n; ++ii) {
        D += p.delay[ii];
        reach = std::max(reach, D + std::max(0.0, p.width[ii]));
    }
    return period <= reach;
}

}  // namespace hsg
