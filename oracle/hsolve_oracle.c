/*
 * hsolve_oracle.c — CPU restatement of MOOSE HSolve (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see hsolve_oracle.h).  Each function cites the
 * reference function whose arithmetic it restates; the operation order is kept
 * so that, built with -ffp-contract=off, results are bit-identical to the
 * reference's.  Pointers into HS/HJ/VMid play the role of the reference's
 * vector iterators.
 */
#include "hsolve_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small growable arrays                                                     */
/* ------------------------------------------------------------------------ */
typedef struct { int* a; int n, cap; } ivec;
static void iv_push(ivec* v, int x) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 8; v->a = (int*)realloc(v->a, sizeof(int) * v->cap); }
    v->a[v->n++] = x;
}
static void iv_free(ivec* v) { free(v->a); v->a = NULL; v->n = v->cap = 0; }

/* ------------------------------------------------------------------------ */
/* LookupTable — hsolve/RateLookup.cpp:15-103                                */
/* ------------------------------------------------------------------------ */
typedef struct {
    double  min, max, dx;
    unsigned nPts, nColumns;
    double* table;
} or_table;

typedef struct { double* row; double fraction; } or_row;

/* LookupTable::LookupTable — RateLookup.cpp:15-31 */
static void table_init(or_table* t, double min, double max, unsigned nDivs, unsigned nSpecies) {
    t->min = min;
    t->max = max;
    t->nPts = nDivs + 1 + 1;
    t->dx = (max - min) / nDivs;
    t->nColumns = 2 * nSpecies;
    size_t sz = (size_t)t->nPts * t->nColumns;
    t->table = sz ? (double*)calloc(sz, sizeof(double)) : NULL;
}

/* LookupTable::addColumns — RateLookup.cpp:33-55 (C1, C2 hold nPts-1 values) */
static void table_add_columns(or_table* t, int species, const double* C1, const double* C2, unsigned nC) {
    double* it = t->table + 2 * species;
    unsigned ig;
    for (ig = 0; ig < t->nPts - 1; ++ig) {
        it[0] = C1[ig];
        it[1] = C2[ig];
        it += t->nColumns;
    }
    it[0] = C1[nC - 1];
    it[1] = C2[nC - 1];
}

/* LookupTable::row — RateLookup.cpp:63-76 */
static void table_row(const or_table* t, double x, or_row* r) {
    if (x < t->min) x = t->min;
    else if (x > t->max) x = t->max;
    double div = (x - t->min) / t->dx;
    unsigned integer = (unsigned)(div);
    r->fraction = div - integer;
    r->row = t->table ? t->table + (size_t)integer * t->nColumns : NULL;
}

/* LookupTable::lookup — RateLookup.cpp:78-103 */
static void table_lookup(const or_table* t, unsigned column, const or_row* r, double* C1, double* C2) {
    double a, b;
    const double* ap = r->row + column;
    const double* bp = ap + t->nColumns;
    a = *ap;
    b = *bp;
    *C1 = a + (b - a) * r->fraction;
    a = *(ap + 1);
    b = *(bp + 1);
    *C2 = a + (b - a) * r->fraction;
}

double or_lookup_row(double min, double max, unsigned nDivs, double x, unsigned* index) {
    or_table t;
    t.min = min; t.max = max; t.nPts = nDivs + 2; t.dx = (max - min) / nDivs; t.nColumns = 0; t.table = NULL;
    or_row r;
    table_row(&t, x, &r);
    if (x < t.min) x = t.min; else if (x > t.max) x = t.max;
    *index = (unsigned)((x - t.min) / t.dx);
    return r.fraction;
}

/* ------------------------------------------------------------------------ */
/* solver state (mirrors HinesMatrix / HSolvePassive / HSolveActive members) */
/* ------------------------------------------------------------------------ */
typedef struct { double c, CaBasal, factor1, factor2, ceiling, floor; } or_caconc;

typedef struct {
    int    compt;        /* Hines index */
    double threshold, refractT, lastEvent, V;
    int    fired, edgeTriggered;
} or_spikegen;

typedef struct {
    /* HinesMatrix (hsolve/HinesMatrix.h:67-88) */
    unsigned nCompt;
    double   dt;                 /* HinesMatrix::dt_ (setup dt) */
    unsigned nJunc;
    unsigned *jIndex, *jRank;    /* junction_ */
    double  *HS, *HJ, *HJCopy, *VMid;
    unsigned nHJ;
    double **operand;  unsigned nOp;
    double **backOperand; unsigned nBop;
    int      stage;
    /* tree_ (TreeNodeStruct) */
    int     *cstart, *children;  /* children of i: children[cstart[i]..cstart[i+1]) */
    double  *Ra, *Rm, *Cm, *Em, *initVm;
    /* makeJunctions bookkeeping */
    int      nGroups;
    int     *gstart, *gmem;      /* coupled_: sorted member lists */
    int     *groupNumber;        /* per compartment, -1 if none */
    int     *operandBase;        /* per compartment: offset into HJ, -1 */
    /* HSolvePassive */
    double  *CmByDt, *EmByRm, *V;
    int     *hasInject;
    double  *injectVarying, *injectBasal;
} or_hines;

/* A PulseGen (device/PulseGen.cpp) sending to a compartment's injectMsg
 * (ic >= 0) and/or to other PulseGens' input. */
typedef struct {
    int    ic;                        /* Hines index of the target, -1 */
    int    n;                         /* entries of delay/width/level */
    int    trigMode;                  /* trigMode_ */
    int    src;                       /* the PulseGen whose output drives input_, -1 */
    double base, out;                 /* baseLevel_, output_ */
    double input, prevInput, trigTime;/* input_, prevInput_, trigTime_ */
    double delay[HSG_MAX_PULSES], width[HSG_MAX_PULSES], level[HSG_MAX_PULSES];
} or_pulse;

/* An Interpol2D (builtins/Interpol2D.cpp) and an HHChannel2D
 * (biophysics/HHChannel2D.cpp) beside the solver. */
typedef struct {
    unsigned nx, ny;                  /* xdivs+1, ydivs+1 */
    double   xmin, xmax, ymin, ymax, invDx, invDy;
    const double* t;                  /* [ix*ny + iy], owned copy */
} or_tab2d;

typedef struct {
    int    compt;                     /* Hines index */
    int    instant;
    int    tabA[3], tabB[3];
    int    dep[6];
    int    conc1, conc2;              /* oracle pool indices, -1 */
    double Gbar, Ek, pw[3], st[3], mod, Gk;
} or_chan2d;

/* A VClamp (biophysics/VClamp.cpp) sensing a compartment's Vm and sending
 * its current to the compartment's injectMsg. */
typedef struct {
    int    ic;                        /* Hines index */
    int    mode;
    double tau, ti, td, Kp;           /* fields (Kp = gain) */
    double expt, tauByDt, dtByTi, tdByDt;
    double vIn, v1, cmdIn, oldCmdIn, command, e, e1, e2, current;
} or_vclamp;

/* doubleEq (basecode/doubleEq.cpp:16-22) */
static int double_eq(double x, double y) {
    double denom = fabs(x) + fabs(y);
    if (denom < 1e-12) denom = 1e-6;
    return (fabs(x - y) / denom) < 1e-6;
}

/* A Leakage object's per-tick send (biophysics/Leakage.cpp:90-95). */
typedef struct {
    int    ic;                        /* Hines index */
    double Gbar, Ek, mod;
} or_leak;

/* A pending synaptic event (synapse/SynEvent.h). */
typedef struct {
    double time, weight;
} or_event;

/* A SynChan (biophysics/SynChan.cpp) and its SimpleSynHandler
 * (synapse/SimpleSynHandler.cpp): the handler's queue holds events sorted
 * by time, equal times in arrival order. */
typedef struct {
    int    ic;                        /* Hines index */
    int    firstSyn, nSyn;
    double Gbar, Ek, tau1, tau2, mod;
    double xconst1, xconst2, yconst1, yconst2, norm, dt;
    double X, Y, Gk, activation;
    or_event* ev;
    int    nEv, capEv;
} or_synchan;

/* SynChan's integration constants and norm at step dt: setTau1/setTau2
 * and normalizeGbar (biophysics/SynChan.cpp:123-198), as reinit computes
 * them (:229-251). */
static void synchan_constants(or_synchan* c, double dt) {
    c->dt = dt;
    c->xconst1 = c->tau1 * (1.0 - exp(-dt / c->tau1));
    c->xconst2 = exp(-dt / c->tau1);
    if (double_eq(c->tau2, 0.0)) {
        c->yconst1 = 1.0;
        c->yconst2 = 0.0;
    } else {
        c->yconst1 = c->tau2 * (1.0 - exp(-dt / c->tau2));
        c->yconst2 = exp(-dt / c->tau2);
    }
    if (double_eq(c->tau2, 0.0)) {
        c->norm = c->Gbar;
    } else if (double_eq(c->tau1, c->tau2)) {
        c->norm = c->Gbar * exp(1.0) / c->tau1;
    } else {
        double tpeak = c->tau1 * c->tau2 * log(c->tau1 / c->tau2) / (c->tau1 - c->tau2);
        c->norm = c->Gbar * (c->tau1 - c->tau2) /
                  (c->tau1 * c->tau2 * (exp(-tpeak / c->tau1) - exp(-tpeak / c->tau2)));
    }
}

struct or_solver {
    or_hines h;
    double   hsolveDt;           /* HSolve::dt_ (shadows HinesMatrix::dt_) */
    int      caAdvance;
    int     *hinesToDesc, *descToHines;
    /* channels (HSolveActive.h:73-136) */
    int      nChan;
    double  *Gbar, *Xpower, *Ypower, *Zpower, *modulation;
    int     *instant;
    double  *Gk, *Ek;            /* current_ */
    int      nState;
    double  *state;
    int     *chan2state, *chan2compt, *channelCount, *currentBoundary;
    int     *chanToDesc, *descToChan;
    /* calcium */
    int      nPool;
    or_caconc* caConc;
    double  *ca, *caActivation;
    int     *caTarget;           /* pool index or -1 (caTarget_ pointers) */
    int     *caCount;
    int     *poolToDesc, *descToPool;
    /* lookup */
    int      nColumn;
    unsigned *column;
    int      nZ;
    int     *caRowIdx;           /* caRow_: index into caRowCompt, -1 = null */
    int      maxCaCount;
    or_row  *caRowCompt;
    or_table vTable, caTable;
    /* external inputs */
    double  *externalCalcium, *externalCurrent, *prevExtCurr;
    /* spikes */
    int      nSpike;
    or_spikegen* spike;
    int     *spikeToDesc;
    ivec     spikeGenLog;        /* recorded spikes */
    long    *spikeStepLog; int spikeStepCap;
    long     stepCount;          /* steps since reinit */
    /* HHChannel2D objects, desc order */
    int        n2d, nTab2d;
    or_chan2d* ch2d;
    or_tab2d*  tab2d;
    int        valuesSent;       /* the solver has sent Vm/[Ca] at least once */
    /* PulseGens feeding injectMsg, desc order */
    int        nPulse;
    or_pulse*  pulse;
    /* Leakage objects and SynChans sending to the compartments, desc order */
    int         nLeak;
    or_leak*    leak;
    int         nSyn, nSynapse;
    or_synchan* syn;
    double     *synW, *synDelay;      /* synapses, desc order */
    /* VClamps on the compartments, desc order */
    int         nVc;
    or_vclamp*  vc;
};

/* ------------------------------------------------------------------------ */
/* HinesMatrix — hsolve/HinesMatrix.cpp                                     */
/* ------------------------------------------------------------------------ */

/* insertion sort keeps this deterministic; groups are small */
static void sort_ints(int* a, int n) {
    for (int i = 1; i < n; ++i) {
        int x = a[i], j = i - 1;
        while (j >= 0 && a[j] > x) { a[j + 1] = a[j]; --j; }
        a[j + 1] = x;
    }
}

/* makeJunctions — HinesMatrix.cpp:85-133 */
static void hm_make_junctions(or_hines* h) {
    unsigned n = h->nCompt;
    ivec gs = {0}, gm = {0};
    iv_push(&gs, 0);
    for (unsigned i = 0; i < n; ++i) {
        int nc = h->cstart[i + 1] - h->cstart[i];
        const int* c = h->children + h->cstart[i];
        if (nc == 0) continue;
        if (nc == 1) {
            int diff = c[0] - (int)i;
            if (diff == 1 || diff == -1) continue;
        }
        for (int k = 0; k < nc; ++k) iv_push(&gm, c[k]);
        iv_push(&gm, (int)i);
        iv_push(&gs, gm.n);
    }
    h->nGroups = gs.n - 1;
    /* 3.2: sort each group, then groups by first element (stable) */
    for (int g = 0; g < h->nGroups; ++g) sort_ints(gm.a + gs.a[g], gs.a[g + 1] - gs.a[g]);
    int* order = (int*)malloc(sizeof(int) * (h->nGroups + 1));
    for (int g = 0; g < h->nGroups; ++g) order[g] = g;
    for (int i = 1; i < h->nGroups; ++i) {
        int x = order[i], j = i - 1;
        while (j >= 0 && gm.a[gs.a[order[j]]] > gm.a[gs.a[x]]) { order[j + 1] = order[j]; --j; }
        order[j + 1] = x;
    }
    h->gstart = (int*)malloc(sizeof(int) * (h->nGroups + 1));
    h->gmem = (int*)malloc(sizeof(int) * (gm.n ? gm.n : 1));
    int pos = 0;
    for (int g = 0; g < h->nGroups; ++g) {
        h->gstart[g] = pos;
        for (int k = gs.a[order[g]]; k < gs.a[order[g] + 1]; ++k) h->gmem[pos++] = gm.a[k];
    }
    h->gstart[h->nGroups] = pos;
    free(order);
    iv_free(&gs); iv_free(&gm);
    /* 3.3 */
    h->groupNumber = (int*)malloc(sizeof(int) * n);
    for (unsigned i = 0; i < n; ++i) h->groupNumber[i] = -1;
    unsigned nj = 0;
    for (int g = 0; g < h->nGroups; ++g) nj += (unsigned)(h->gstart[g + 1] - h->gstart[g] - 1);
    h->jIndex = (unsigned*)malloc(sizeof(unsigned) * (nj + 1));
    h->jRank = (unsigned*)malloc(sizeof(unsigned) * (nj + 1));
    h->nJunc = 0;
    for (int g = 0; g < h->nGroups; ++g) {
        int size = h->gstart[g + 1] - h->gstart[g];
        for (int c = 0; c < size - 1; ++c) {
            h->jIndex[h->nJunc] = (unsigned)h->gmem[h->gstart[g] + c];
            h->jRank[h->nJunc] = (unsigned)(size - c - 1);
            h->nJunc++;
            h->groupNumber[h->gmem[h->gstart[g] + c]] = g;
        }
    }
    /* sort( junction_ ) by index (stable insertion sort) */
    for (unsigned i = 1; i < h->nJunc; ++i) {
        unsigned xi = h->jIndex[i], xr = h->jRank[i];
        int j = (int)i - 1;
        while (j >= 0 && h->jIndex[j] > xi) { h->jIndex[j + 1] = h->jIndex[j]; h->jRank[j + 1] = h->jRank[j]; --j; }
        h->jIndex[j + 1] = xi; h->jRank[j + 1] = xr;
    }
}

/* makeMatrix — HinesMatrix.cpp:136-222 */
static void hm_make_matrix(or_hines* h, const double* Ga) {
    unsigned n = h->nCompt;
    h->HS = (double*)calloc(4 * (size_t)n, sizeof(double));
    for (unsigned i = 0; i < n; ++i)
        h->HS[4 * i + 2] = h->Cm[i] / (h->dt / 2.0) + 1.0 / h->Rm[i];

    double gi, gj, gij;
    unsigned jn = 0;
    for (unsigned i = 0; i + 1 < n; ++i) {
        if (h->nJunc != 0 && jn < h->nJunc && i == h->jIndex[jn]) { ++jn; continue; }
        gi = Ga[i];
        gj = Ga[i + 1];
        gij = gi * gj / (gi + gj);
        h->HS[4 * i + 1] = -gij;
        h->HS[4 * i + 2] += gij;
        h->HS[4 * i + 6] += gij;
    }
    for (int g = 0; g < h->nGroups; ++g) {
        double gsum = 0.0;
        for (int k = h->gstart[g]; k < h->gstart[g + 1]; ++k) gsum += Ga[h->gmem[k]];
        for (int k = h->gstart[g]; k < h->gstart[g + 1]; ++k) {
            gi = Ga[h->gmem[k]];
            h->HS[4 * h->gmem[k] + 2] += gi * (1.0 - gi / gsum);
        }
    }
    unsigned size = 0;
    for (int g = 0; g < h->nGroups; ++g) {
        unsigned rank = (unsigned)(h->gstart[g + 1] - h->gstart[g] - 1);
        size += rank * (rank + 1);
    }
    h->nHJ = size;
    h->HJ = (double*)malloc(sizeof(double) * (size ? size : 1));
    h->operandBase = (int*)malloc(sizeof(int) * n);
    for (unsigned i = 0; i < n; ++i) h->operandBase[i] = -1;
    unsigned pos = 0;
    for (int g = 0; g < h->nGroups; ++g) {
        double gsum = 0.0;
        for (int k = h->gstart[g]; k < h->gstart[g + 1]; ++k) gsum += Ga[h->gmem[k]];
        for (int a = h->gstart[g]; a < h->gstart[g + 1] - 1; ++a) {
            int base = (int)pos;
            for (int b = a + 1; b < h->gstart[g + 1]; ++b) {
                gij = Ga[h->gmem[a]] * Ga[h->gmem[b]] / gsum;
                h->HJ[pos++] = -gij;
                h->HJ[pos++] = -gij;
            }
            h->operandBase[h->gmem[a]] = base;
        }
    }
    for (unsigned i = 0; i < n; ++i) h->HS[4 * i] = h->HS[4 * i + 2];
    h->HJCopy = (double*)malloc(sizeof(double) * (size ? size : 1));
    memcpy(h->HJCopy, h->HJ, sizeof(double) * size);
}

/* makeOperands — HinesMatrix.cpp:225-365 */
static void hm_make_operands(or_hines* h) {
    unsigned n = h->nCompt;
    h->VMid = (double*)calloc(n ? n : 1, sizeof(double));
    unsigned nop = 0, nbop = 0;
    for (unsigned j = 0; j < h->nJunc; ++j) {
        unsigned r = h->jRank[j];
        nop += (r == 1) ? 3 : (r == 2) ? 5 : 3 * r * (r + 1);
        if (r >= 3) nbop += 2 * r;
    }
    h->operand = (double**)malloc(sizeof(double*) * (nop ? nop : 1));
    h->backOperand = (double**)malloc(sizeof(double*) * (nbop ? nbop : 1));
    double** op = h->operand;
    for (unsigned jn = 0; jn < h->nJunc; ++jn) {
        unsigned index = h->jIndex[jn], rank = h->jRank[jn];
        double* base = h->HJ + h->operandBase[index];
        int g = h->groupNumber[index];
        const int* group = h->gmem + h->gstart[g];
        int gsize = h->gstart[g + 1] - h->gstart[g];
        unsigned far;
        if (rank == 1) {
            *op++ = base;
            far = (unsigned)group[gsize - 1];
            *op++ = h->HS + 4 * far;
            *op++ = h->VMid + far;
        } else if (rank == 2) {
            *op++ = base;
            far = (unsigned)group[gsize - 2];
            *op++ = h->HS + 4 * far;
            *op++ = h->VMid + far;
            far = (unsigned)group[gsize - 1];
            *op++ = h->HS + 4 * far;
            *op++ = h->VMid + far;
        } else {
            int start = gsize - (int)rank;
            for (unsigned j = 0; j < rank; ++j) {
                far = (unsigned)group[start + j];
                *op++ = h->HS + 4 * far;
                *op++ = base + 2 * j;
                *op++ = base + 2 * j + 1;
                *op++ = h->HS + 4 * far + 3;
                *op++ = h->HS + 4 * index + 3;
                *op++ = base + 2 * j + 1;
            }
            double *left, *above, *target;
            left = base + 1;
            target = base + 2 * rank;
            for (unsigned i = 1; i < rank; ++i) {
                above = base + 2 * i;
                for (unsigned j = 0; j < rank - i; ++j) {
                    *op++ = target; *op++ = above; *op++ = left;
                    above += 2;
                    target += 2;
                }
                left += 2;
            }
            target = base + 2 * rank + 1;
            above = base;
            for (unsigned i = 1; i < rank; ++i) {
                left = base + 2 * i + 1;
                for (unsigned j = 0; j < rank - i; ++j) {
                    *op++ = target; *op++ = above; *op++ = left;
                    if (i == rank - 1 && j == rank - i - 1) continue;
                    target += 2;
                    left += 2;
                }
                above += 2;
            }
        }
    }
    h->nOp = (unsigned)(op - h->operand);
    double** bop = h->backOperand;
    for (unsigned jn = 0; jn < h->nJunc; ++jn) {
        if (h->jRank[jn] < 3) continue;
        unsigned index = h->jIndex[jn], rank = h->jRank[jn];
        double* base = h->HJ + h->operandBase[index];
        int g = h->groupNumber[index];
        const int* group = h->gmem + h->gstart[g];
        int gsize = h->gstart[g + 1] - h->gstart[g];
        unsigned start = (unsigned)gsize - rank;
        for (unsigned j = 0; j < rank; ++j) {
            unsigned far = (unsigned)group[start + j];
            *bop++ = base + 2 * j;
            *bop++ = h->VMid + far;
        }
    }
    h->nBop = (unsigned)(bop - h->backOperand);
}

/* HinesMatrix::setup — HinesMatrix.cpp:25-52 */
static void hm_setup(or_hines* h) {
    unsigned n = h->nCompt;
    double* Ga = (double*)malloc(sizeof(double) * (n ? n : 1));
    for (unsigned i = 0; i < n; ++i) Ga[i] = 2.0 / h->Ra[i];
    hm_make_junctions(h);
    hm_make_matrix(h, Ga);
    hm_make_operands(h);
    h->stage = 0;
    free(Ga);
}

/* HinesMatrix::getA — HinesMatrix.cpp:376-431 */
static double hm_getA(const or_hines* h, unsigned row, unsigned col) {
    if ((h->stage == 1 || h->stage == 2) && row > col) return 0.0;
    if (row >= h->nCompt || col >= h->nCompt) return 0.0;
    if (row == col) return h->HS[4 * row];
    unsigned smaller = row < col ? row : col;
    unsigned bigger = row > col ? row : col;
    if (h->groupNumber[smaller] < 0) {
        if (bigger - smaller == 1) return h->HS[4 * smaller + 1];
        return 0.0;
    }
    int g = h->groupNumber[smaller];
    const int* group = h->gmem + h->gstart[g];
    int size = h->gstart[g + 1] - h->gstart[g];
    int ps = -1, pb = -1;
    for (int k = 0; k < size; ++k) {
        if (group[k] == (int)smaller) ps = k;
        if (group[k] == (int)bigger) pb = k;
    }
    if (pb < 0) return 0.0;
    unsigned location = 0;
    for (int i = 0; i < g; ++i) {
        unsigned sz = (unsigned)(h->gstart[i + 1] - h->gstart[i]);
        location += sz * (sz - 1);
    }
    unsigned smallRank = (unsigned)(size - ps - 1);
    unsigned bigRank = (unsigned)(size - pb - 1);
    location += (unsigned)size * (unsigned)(size - 1) - smallRank * (smallRank + 1);
    location += 2 * (smallRank - bigRank - 1);
    return row == smaller ? h->HJ[location] : h->HJ[location + 1];
}

/* HSolvePassive::updateMatrix — HSolvePassive.cpp:211-244 */
static void hp_update_matrix(or_hines* h) {
    if (h->nHJ) memcpy(h->HJ, h->HJCopy, sizeof(double) * h->nHJ);
    for (unsigned i = 0; i < h->nCompt; ++i) {
        double* ihs = h->HS + 4 * i;
        ihs[0] = ihs[2];
        ihs[3] = h->V[i] * h->CmByDt[i] + h->EmByRm[i];
    }
    for (unsigned i = 0; i < h->nCompt; ++i) {
        if (!h->hasInject[i]) continue;
        h->HS[4 * i + 3] += h->injectVarying[i] + h->injectBasal[i];
        h->injectVarying[i] = 0.0;
    }
    h->stage = 0;
}

/* HSolvePassive::forwardEliminate — HSolvePassive.cpp:246-323 */
static void hp_forward(or_hines* h) {
    unsigned ic = 0;
    double* ihs = h->HS;
    double** iop = h->operand;
    double pivot, division;
    for (unsigned jn = 0; jn < h->nJunc; ++jn) {
        unsigned index = h->jIndex[jn], rank = h->jRank[jn];
        while (ic < index) {
            *(ihs + 4) -= *(ihs + 1) / *ihs * *(ihs + 1);
            *(ihs + 7) -= *(ihs + 1) / *ihs * *(ihs + 3);
            ++ic, ihs += 4;
        }
        pivot = *ihs;
        if (rank == 1) {
            double* j = *iop;
            double* s = *(iop + 1);
            division = *(j + 1) / pivot;
            *(s) -= division * *j;
            *(s + 3) -= division * *(ihs + 3);
            iop += 3;
        } else if (rank == 2) {
            double* j = *iop;
            double* s;
            s = *(iop + 1);
            division = *(j + 1) / pivot;
            *(s) -= division * *j;
            *(j + 4) -= division * *(j + 2);
            *(s + 3) -= division * *(ihs + 3);
            s = *(iop + 3);
            division = *(j + 3) / pivot;
            *(j + 5) -= division * *j;
            *(s) -= division * *(j + 2);
            *(s + 3) -= division * *(ihs + 3);
            iop += 5;
        } else {
            double** end = iop + 3 * rank * (rank + 1);
            for (; iop < end; iop += 3)
                **iop -= **(iop + 2) / pivot * **(iop + 1);
        }
        ++ic, ihs += 4;
    }
    while (ic + 1 < h->nCompt) {
        *(ihs + 4) -= *(ihs + 1) / *ihs * *(ihs + 1);
        *(ihs + 7) -= *(ihs + 1) / *ihs * *(ihs + 3);
        ++ic, ihs += 4;
    }
    h->stage = 1;
}

/* HSolvePassive::backwardSubstitute — HSolvePassive.cpp:325-401.
 * The reference walks HS_ with a reverse iterator; here rs(k) = HS[4n-1-k]. */
static void hp_backward(or_hines* h) {
    int n = (int)h->nCompt;
    int ic = n - 1;
    double* HS = h->HS;
    double* V = h->V;
    double* VMid = h->VMid;
    long r = 4L * n - 1;                 /* reverse-iterator position into HS */
#define RHS(k) HS[r - (k)]
    long iop = (long)h->nOp - 1;         /* reverse iterator into operand_   */
    long ibop = (long)h->nBop - 1;       /* reverse iterator into backOperand_ */
    int iv = n - 1;

    VMid[iv] = RHS(0) / RHS(3);
    V[iv] = 2 * VMid[iv] - V[iv];
    --ic, --iv, r -= 4;

    for (int jn = (int)h->nJunc - 1; jn >= 0; --jn) {
        int index = (int)h->jIndex[jn];
        int rank = (int)h->jRank[jn];
        while (ic > index) {
            VMid[iv] = (RHS(0) - RHS(2) * VMid[iv + 1]) / RHS(3);
            V[iv] = 2 * VMid[iv] - V[iv];
            --ic, --iv, r -= 4;
        }
        if (rank == 1) {
            /* iop[0] = VMid+far, iop[2] = base */
            VMid[iv] = (RHS(0) - *h->operand[iop] * *h->operand[iop - 2]) / RHS(3);
            iop -= 3;
        } else if (rank == 2) {
            double* v0 = h->operand[iop];
            double* v1 = h->operand[iop - 2];
            double* j = h->operand[iop - 4];
            VMid[iv] = (RHS(0) - *v0 * *(j + 2) - *v1 * *j) / RHS(3);
            iop -= 5;
        } else {
            VMid[iv] = RHS(0);
            for (int i = 0; i < rank; ++i) {
                VMid[iv] -= *h->backOperand[ibop] * *h->backOperand[ibop - 1];
                ibop -= 2;
            }
            VMid[iv] /= RHS(3);
            iop -= 3 * rank * (rank + 1);
        }
        V[iv] = 2 * VMid[iv] - V[iv];
        --ic, --iv, r -= 4;
    }
    while (ic >= 0) {
        VMid[iv] = (RHS(0) - RHS(2) * VMid[iv + 1]) / RHS(3);
        V[iv] = 2 * VMid[iv] - V[iv];
        --ic, --iv, r -= 4;
    }
#undef RHS
    h->stage = 2;
}

static void hm_free(or_hines* h) {
    free(h->jIndex); free(h->jRank); free(h->HS); free(h->HJ); free(h->HJCopy); free(h->VMid);
    free(h->operand); free(h->backOperand); free(h->cstart); free(h->children);
    free(h->Ra); free(h->Rm); free(h->Cm); free(h->Em); free(h->initVm);
    free(h->gstart); free(h->gmem); free(h->groupNumber); free(h->operandBase);
    free(h->CmByDt); free(h->EmByRm); free(h->V); free(h->hasInject);
    free(h->injectVarying); free(h->injectBasal);
}

/* ------------------------------------------------------------------------ */
/* walkTree — hsolve/HSolvePassive.cpp:45-98 with HSolveUtils::adjacent      */
/* (hsolve/HSolveUtils.cpp:26-45): axial targets, then raxial targets.       */
/* ------------------------------------------------------------------------ */
/* Per-compartment message lists, kept in message order. */
typedef struct {
    int  n;
    int *axStart, *axList;     /* axial targets (children) of each compartment */
    int *raStart, *raList;     /* raxial targets (parents) of each compartment */
} or_adj;

static void by_key(int nkeys, int nitems, const int* key, int** start, int** list) {
    int* st = (int*)calloc((size_t)nkeys + 1, sizeof(int));
    for (int k = 0; k < nitems; ++k) if (key[k] >= 0 && key[k] < nkeys) st[key[k] + 1]++;
    for (int i = 0; i < nkeys; ++i) st[i + 1] += st[i];
    int* fill = (int*)malloc(sizeof(int) * ((size_t)nkeys + 1));
    memcpy(fill, st, sizeof(int) * (size_t)nkeys);
    int* li = (int*)malloc(sizeof(int) * ((size_t)nitems + 1));
    for (int k = 0; k < nitems; ++k) if (key[k] >= 0 && key[k] < nkeys) li[fill[key[k]]++] = k;
    free(fill);
    *start = st;
    *list = li;
}

static void adj_build(const hsg_cell_desc* d, or_adj* a) {
    int n = d->n_compartments, m = d->n_axial;
    int* par = (int*)malloc(sizeof(int) * ((size_t)m + 1));
    int* chi = (int*)malloc(sizeof(int) * ((size_t)m + 1));
    for (int k = 0; k < m; ++k) { par[k] = d->axial[2 * k]; chi[k] = d->axial[2 * k + 1]; }
    a->n = n;
    by_key(n, m, par, &a->axStart, &a->axList);   /* message k indexed by its parent */
    by_key(n, m, chi, &a->raStart, &a->raList);   /* message k indexed by its child  */
    for (int i = 0; i < a->axStart[n]; ++i) a->axList[i] = chi[a->axList[i]];
    for (int i = 0; i < a->raStart[n]; ++i) a->raList[i] = par[a->raList[i]];
    free(par); free(chi);
}

static void adj_free(or_adj* a) { free(a->axStart); free(a->axList); free(a->raStart); free(a->raList); }

/* HSolveUtils::adjacent(compartment, exclude, ret) */
static void adjacent_of(const or_adj* a, int c, int exclude, ivec* out) {
    for (int k = a->axStart[c]; k < a->axStart[c + 1]; ++k)
        if (a->axList[k] != exclude) iv_push(out, a->axList[k]);
    for (int k = a->raStart[c]; k < a->raStart[c + 1]; ++k)
        if (a->raList[k] != exclude) iv_push(out, a->raList[k]);
}

static int walk_tree(const hsg_cell_desc* d, const or_adj* A, int* order /* n */, int* nout) {
    int n = d->n_compartments;
    int seed = d->seed;
    ivec adj = {0};
    adjacent_of(A, seed, -2, &adj);
    if (adj.n > 1) {
        int guard = 0;
        while (adj.n) {
            int previous = seed;
            seed = adj.a[0];
            adj.n = 0;
            adjacent_of(A, seed, previous, &adj);
            if (++guard > n + 1) { iv_free(&adj); return HSG_E_MODEL; }
        }
    }
    iv_free(&adj);
    /* depth-first search with a stack of adjacency lists */
    ivec* cstack = (ivec*)calloc((size_t)n + 2, sizeof(ivec));
    int depth = 1;
    iv_push(&cstack[0], seed);
    int above = -2;
    int count = 0;
    while (depth > 0) {
        ivec* top = &cstack[depth - 1];
        if (top->n == 0) {
            iv_free(top);
            --depth;
            if (depth > 0) cstack[depth - 1].n--;
        } else {
            if (depth > 1) above = cstack[depth - 2].a[cstack[depth - 2].n - 1];
            int current = top->a[top->n - 1];
            if (count >= n || depth >= n + 1) { /* a loop in the morphology */
                for (int k = 0; k < n + 2; ++k) iv_free(&cstack[k]);
                free(cstack);
                return HSG_E_MODEL;
            }
            order[count++] = current;
            cstack[depth].n = 0;
            adjacent_of(A, current, above, &cstack[depth]);
            ++depth;
        }
    }
    free(cstack);
    /* reverse */
    for (int i = 0; i < count / 2; ++i) { int t = order[i]; order[i] = order[count - 1 - i]; order[count - 1 - i] = t; }
    *nout = count;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* HHGate::lookupBoth with interpolation on — biophysics/HHGate.cpp:258-281 */
/* ------------------------------------------------------------------------ */
static void gate_lookup_both(const hsg_gate_desc* g, double v, double* A, double* B) {
    double xmin = g->min, xmax = g->max;
    double invDx = (double)g->divs / (xmax - xmin);
    if (v <= xmin) { *A = g->tableA[0]; *B = g->tableB[0]; }
    else if (v >= xmax) { *A = g->tableA[g->divs]; *B = g->tableB[g->divs]; }
    else {
        unsigned index = (unsigned)((v - xmin) * invDx);
        double frac = (v - xmin - index / invDx) * invDx;
        *A = g->tableA[index] * (1 - frac) + g->tableA[index + 1] * frac;
        *B = g->tableB[index] * (1 - frac) + g->tableB[index + 1] * frac;
    }
}

/* HSolveUtils::rates — hsolve/HSolveUtils.cpp:205-252 */
static void gate_rates(const hsg_gate_desc* g, double gmin, double gmax, unsigned gdivs,
                       double* A, double* B) {
    double dx = (gmax - gmin) / gdivs;
    if (gmin == g->min && gmax == g->max && gdivs == g->divs) {
        memcpy(A, g->tableA, sizeof(double) * (g->divs + 1));
        memcpy(B, g->tableB, sizeof(double) * (g->divs + 1));
        return;
    }
    for (unsigned i = 0; i < gdivs + 1; ++i) gate_lookup_both(g, gmin + dx * i, &A[i], &B[i]);
}

/* ------------------------------------------------------------------------ */
/* setup — HSolve::setup -> HSolveActive::setup                              */
/* ------------------------------------------------------------------------ */
static int cmp_gate_id(const void* a, const void* b, const hsg_cell_desc* d) {
    int64_t x = d->gates[*(const int*)a].id, y = d->gates[*(const int*)b].id;
    return (x > y) - (x < y);
}

/* sorted unique list of gate indices by id (std::set<Id>) */
static int unique_by_id(const hsg_cell_desc* d, int* list, int n) {
    for (int i = 1; i < n; ++i) {
        int x = list[i], j = i - 1;
        while (j >= 0 && cmp_gate_id(&list[j], &x, d) > 0) { list[j + 1] = list[j]; --j; }
        list[j + 1] = x;
    }
    int m = 0;
    for (int i = 0; i < n; ++i)
        if (m == 0 || d->gates[list[m - 1]].id != d->gates[list[i]].id) list[m++] = list[i];
    return m;
}

int or_create(const hsg_cell_desc* d, double dt, int caAdvance, or_solver** out) {
    if (!d || !out || d->n_compartments <= 0 || !(dt > 0.0)) return HSG_E_ARG;
    if (d->seed < 0 || d->seed >= d->n_compartments) return HSG_E_ARG;
    or_solver* s = (or_solver*)calloc(1, sizeof(or_solver));
    or_hines* h = &s->h;
    int nd = d->n_compartments;
    s->hinesToDesc = (int*)malloc(sizeof(int) * nd);
    int n = 0;
    or_adj A;
    adj_build(d, &A);
    int rc = walk_tree(d, &A, s->hinesToDesc, &n);
    if (rc) { adj_free(&A); free(s->hinesToDesc); free(s); return rc; }
    /* per-compartment lists of channels, leaks, spikegens in message order */
    int *chStart, *chList, *lkStart, *lkList, *sgStart, *sgList;
    {
        int* key = (int*)malloc(sizeof(int) * ((size_t)d->n_channels + d->n_leaks + d->n_spikegens + 1));
        for (int k = 0; k < d->n_channels; ++k) key[k] = d->channels[k].compartment;
        by_key(nd, d->n_channels, key, &chStart, &chList);
        for (int k = 0; k < d->n_leaks; ++k) key[k] = d->leaks[k].compartment;
        by_key(nd, d->n_leaks, key, &lkStart, &lkList);
        for (int k = 0; k < d->n_spikegens; ++k) key[k] = d->spikegens[k].compartment;
        by_key(nd, d->n_spikegens, key, &sgStart, &sgList);
        free(key);
    }
    s->descToHines = (int*)malloc(sizeof(int) * nd);
    for (int i = 0; i < nd; ++i) s->descToHines[i] = -1;
    for (int i = 0; i < n; ++i) s->descToHines[s->hinesToDesc[i]] = i;
    h->nCompt = (unsigned)n;
    h->dt = dt;
    s->hsolveDt = dt;
    s->caAdvance = caAdvance;

    /* HSolvePassive::initialize — HSolvePassive.cpp:100-171 */
    h->V = (double*)malloc(sizeof(double) * n);
    h->CmByDt = (double*)malloc(sizeof(double) * n);
    h->EmByRm = (double*)malloc(sizeof(double) * n);
    h->hasInject = (int*)calloc(n, sizeof(int));
    h->injectVarying = (double*)calloc(n, sizeof(double));
    h->injectBasal = (double*)calloc(n, sizeof(double));
    for (int ic = 0; ic < n; ++ic) {
        const hsg_compartment_desc* c = &d->compartments[s->hinesToDesc[ic]];
        h->V[ic] = c->Vm;
        double EmGmThev = c->Em / c->Rm;
        for (int q = lkStart[s->hinesToDesc[ic]]; q < lkStart[s->hinesToDesc[ic] + 1]; ++q)
            EmGmThev += d->leaks[lkList[q]].Ek * d->leaks[lkList[q]].Gk;
        h->CmByDt[ic] = 2.0 * c->Cm / dt;
        h->EmByRm[ic] = EmGmThev;
        if (c->inject != 0.0) { h->hasInject[ic] = 1; h->injectVarying[ic] = 0.0; h->injectBasal[ic] = c->inject; }
    }
    /* HSolvePassive::storeTree — HSolvePassive.cpp:173-205 */
    h->cstart = (int*)malloc(sizeof(int) * (n + 1));
    h->children = (int*)malloc(sizeof(int) * (d->n_axial + 1));
    h->Ra = (double*)malloc(sizeof(double) * n); h->Rm = (double*)malloc(sizeof(double) * n);
    h->Cm = (double*)malloc(sizeof(double) * n); h->Em = (double*)malloc(sizeof(double) * n);
    h->initVm = (double*)malloc(sizeof(double) * n);
    int pos = 0;
    for (int ic = 0; ic < n; ++ic) {
        int desc = s->hinesToDesc[ic];
        h->cstart[ic] = pos;
        for (int q = A.axStart[desc]; q < A.axStart[desc + 1]; ++q) {
            int ch = s->descToHines[A.axList[q]];
            if (ch < 0) { rc = HSG_E_MODEL; break; }
            h->children[pos++] = ch;
        }
        const hsg_compartment_desc* c = &d->compartments[desc];
        h->Ra[ic] = c->Ra; h->Rm[ic] = c->Rm; h->Cm[ic] = c->Cm; h->Em[ic] = c->Em; h->initVm[ic] = c->initVm;
    }
    h->cstart[n] = pos;
    adj_free(&A);
    free(lkStart); free(lkList);
    if (rc) { free(chStart); free(chList); free(sgStart); free(sgList); or_destroy(s); return rc; }
    hm_setup(h);

    /* HSolveActive::readHHChannels — HSolveActiveSetup.cpp:162-233 */
    s->channelCount = (int*)calloc(n, sizeof(int));
    s->currentBoundary = (int*)calloc(n, sizeof(int));
    int nch = d->n_channels;
    s->Gbar = (double*)malloc(sizeof(double) * (nch + 1));
    s->Xpower = (double*)malloc(sizeof(double) * (nch + 1));
    s->Ypower = (double*)malloc(sizeof(double) * (nch + 1));
    s->Zpower = (double*)malloc(sizeof(double) * (nch + 1));
    s->modulation = (double*)malloc(sizeof(double) * (nch + 1));
    s->instant = (int*)malloc(sizeof(int) * (nch + 1));
    s->Gk = (double*)calloc(nch + 1, sizeof(double));
    s->Ek = (double*)malloc(sizeof(double) * (nch + 1));
    s->state = (double*)malloc(sizeof(double) * (3 * nch + 1));
    s->chan2state = (int*)malloc(sizeof(int) * (nch + 1));
    s->chan2compt = (int*)malloc(sizeof(int) * (nch + 1));
    s->chanToDesc = (int*)malloc(sizeof(int) * (nch + 1));
    s->descToChan = (int*)malloc(sizeof(int) * (nch + 1));
    for (int k = 0; k < nch; ++k) s->descToChan[k] = -1;
    int nc = 0, nstate = 0;
    for (int ic = 0; ic < n; ++ic) {
        int desc = s->hinesToDesc[ic];
        int cnt = 0;
        for (int q = chStart[desc]; q < chStart[desc + 1]; ++q) {
            int k = chList[q];
            const hsg_channel_desc* c = &d->channels[k];
            ++cnt;
            s->Ek[nc] = c->Ek;
            s->Gbar[nc] = c->Gbar;
            s->Xpower[nc] = c->Xpower; s->Ypower[nc] = c->Ypower; s->Zpower[nc] = c->Zpower;
            s->instant[nc] = c->instant;
            s->modulation[nc] = c->modulation;
            s->chan2state[nc] = nstate;
            if (c->Xpower > 0.0) s->state[nstate++] = c->X;
            if (c->Ypower > 0.0) s->state[nstate++] = c->Y;
            if (c->Zpower > 0.0) s->state[nstate++] = c->Z;
            s->chan2compt[nc] = ic;
            s->chanToDesc[nc] = k;
            s->descToChan[k] = nc;
            ++nc;
        }
        s->channelCount[ic] = cnt;
    }
    s->nChan = nc;
    s->nState = nstate;
    free(chStart); free(chList);
    {
        int cum = 0;
        for (int ic = 0; ic < n; ++ic) { cum += s->channelCount[ic]; s->currentBoundary[ic] = cum; }
    }

    /* HSolveActive::readGates — HSolveActiveSetup.cpp:235-248 */
    int* gateId = (int*)malloc(sizeof(int) * (3 * nc + 1));   /* desc gate index */
    int* gCaDepend = (int*)calloc(3 * nc + 1, sizeof(int));
    int ngate = 0;
    for (int c = 0; c < nc; ++c) {
        const hsg_channel_desc* ch = &d->channels[s->chanToDesc[c]];
        double pw[3] = {ch->Xpower, ch->Ypower, ch->Zpower};
        for (int gI = 0; gI < 3; ++gI)
            if (pw[gI] > 0.0) {
                if (ch->gate[gI] < 0 || ch->gate[gI] >= d->n_gates) { free(gateId); free(gCaDepend); or_destroy(s); return HSG_E_MODEL; }
                gateId[ngate] = ch->gate[gI];
                gCaDepend[ngate] = 0;
                ++ngate;
            }
        if (ch->useConcentration && ngate > 0) gCaDepend[ngate - 1] = 1;
    }

    /* HSolveActive::readCalcium — HSolveActiveSetup.cpp:250-337 */
    s->caCount = (int*)calloc(n, sizeof(int));
    int* caTargetIndex = (int*)malloc(sizeof(int) * (nc + 1));
    int* caDependIndex = (int*)malloc(sizeof(int) * (nc + 1));
    s->externalCalcium = (double*)calloc(nc + 1, sizeof(double));
    int npd = d->n_pools;
    int* caConcIndex = (int*)malloc(sizeof(int) * (npd + 1));   /* map<Id,int> */
    for (int p = 0; p < npd; ++p) caConcIndex[p] = -1;
    s->caConc = (or_caconc*)calloc(npd + 1, sizeof(or_caconc));
    s->poolToDesc = (int*)malloc(sizeof(int) * (npd + 1));
    s->descToPool = (int*)malloc(sizeof(int) * (npd + 1));
    for (int p = 0; p < npd; ++p) s->descToPool[p] = -1;
    int npool = 0;
    {
        int ichan = 0;
        for (int ic = 0; ic < n; ++ic) {
            int chanBoundary = ichan + s->channelCount[ic];
            int nCa = npool;
            for (; ichan < chanBoundary; ++ichan) {
                const hsg_channel_desc* ch = &d->channels[s->chanToDesc[ichan]];
                int caConcId[2], ncid = 0;
                int nTarget = 0, nDepend = 0;
                if (ch->caTarget >= 0) { caConcId[ncid++] = ch->caTarget; nTarget = 1; }
                if (nTarget == 0) caTargetIndex[ichan] = -1;
                if (ch->caDepend >= 0) { caConcId[ncid++] = ch->caDepend; nDepend = 1; }
                if (nDepend == 0) caDependIndex[ichan] = -1;
                s->externalCalcium[ichan] = 0;
                for (int q = 0; q < ncid; ++q) {
                    int p = caConcId[q];
                    if (p >= npd) { rc = HSG_E_MODEL; break; }
                    if (caConcIndex[p] < 0) {
                        caConcIndex[p] = s->caCount[ic];
                        ++s->caCount[ic];
                        const hsg_caconc_desc* pc = &d->pools[p];
                        or_caconc* cc = &s->caConc[npool];
                        /* CaConcStruct ctor (HSolveStruct.cpp:103-118): setCa
                         * runs before CaBasal_ is set; reinit resets c_. */
                        cc->CaBasal = 0.0;
                        cc->c = pc->Ca - cc->CaBasal;
                        cc->c += cc->CaBasal - pc->CaBasal;
                        cc->CaBasal = pc->CaBasal;
                        cc->factor1 = 4.0 / (2.0 + dt / pc->tau) - 1.0;
                        cc->factor2 = 2.0 * pc->B * dt / (2.0 + dt / pc->tau);
                        cc->ceiling = pc->ceiling;
                        cc->floor = pc->floor;
                        s->poolToDesc[npool] = p;
                        s->descToPool[p] = npool;
                        ++npool;
                    }
                }
                if (nTarget != 0) caTargetIndex[ichan] = caConcIndex[caConcId[0]] + nCa;
                if (nDepend != 0) caDependIndex[ichan] = caConcIndex[caConcId[ncid - 1]];
            }
        }
    }
    s->nPool = npool;
    s->ca = (double*)calloc(npool + 1, sizeof(double));
    s->caActivation = (double*)calloc(npool + 1, sizeof(double));
    s->caTarget = (int*)malloc(sizeof(int) * (nc + 1));
    for (int c = 0; c < nc; ++c) s->caTarget[c] = caTargetIndex[c];
    free(caConcIndex);
    free(caTargetIndex);
    if (rc) { free(gateId); free(gCaDepend); free(caDependIndex); or_destroy(s); return rc; }

    /* HSolveActive::createLookupTables — HSolveActiveSetup.cpp:339-526 */
    {
        int* caGate = (int*)malloc(sizeof(int) * (ngate + 1));
        int* vGate = (int*)malloc(sizeof(int) * (ngate + 1));
        int nca = 0, nv = 0;
        for (int ig = 0; ig < ngate; ++ig) {
            if (gCaDepend[ig]) caGate[nca++] = gateId[ig];
            else vGate[nv++] = gateId[ig];
        }
        nca = unique_by_id(d, caGate, nca);
        nv = unique_by_id(d, vGate, nv);
        /* gateSpecies: ca ranks first, then v ranks overwrite (map assignment) */
        int* species = (int*)malloc(sizeof(int) * (d->n_gates + 1));
        for (int g = 0; g < d->n_gates; ++g) species[g] = -1;
        for (int ig = 0; ig < nca; ++ig)
            for (int g = 0; g < d->n_gates; ++g)
                if (d->gates[g].id == d->gates[caGate[ig]].id) species[g] = ig;
        for (int ig = 0; ig < nv; ++ig)
            for (int g = 0; g < d->n_gates; ++g)
                if (d->gates[g].id == d->gates[vGate[ig]].id) species[g] = ig;

        double vMin = DBL_MAX, vMax = DBL_MIN, vDx = DBL_MAX;
        double caMin = DBL_MAX, caMax = DBL_MIN, caDx = DBL_MAX;
        for (int ig = 0; ig < nca; ++ig) {
            const hsg_gate_desc* g = &d->gates[caGate[ig]];
            double dx = (g->max - g->min) / g->divs;
            if (g->min < caMin) caMin = g->min;
            if (g->max > caMax) caMax = g->max;
            if (dx < caDx) caDx = dx;
        }
        double caDivD = (caMax - caMin) / caDx;
        int caDiv = (int)(caDivD + 0.5);
        for (int ig = 0; ig < nv; ++ig) {
            const hsg_gate_desc* g = &d->gates[vGate[ig]];
            double dx = (g->max - g->min) / g->divs;
            if (g->min < vMin) vMin = g->min;
            if (g->max > vMax) vMax = g->max;
            if (dx < vDx) vDx = dx;
        }
        double vDivD = (vMax - vMin) / vDx;
        int vDiv = (int)(vDivD + 0.5);
        table_init(&s->caTable, caMin, caMax, (unsigned)caDiv, (unsigned)nca);
        table_init(&s->vTable, vMin, vMax, (unsigned)vDiv, (unsigned)nv);
        unsigned gsz = (unsigned)caDiv + 1;
        double* A = (double*)malloc(sizeof(double) * ((size_t)(vDiv > caDiv ? vDiv : caDiv) + 2));
        double* B = (double*)malloc(sizeof(double) * ((size_t)(vDiv > caDiv ? vDiv : caDiv) + 2));
        for (int ig = 0; ig < nca; ++ig) {
            gate_rates(&d->gates[caGate[ig]], caMin, caMax, (unsigned)caDiv, A, B);
            table_add_columns(&s->caTable, ig, A, B, gsz);
        }
        gsz = (unsigned)vDiv + 1;
        for (int ig = 0; ig < nv; ++ig) {
            gate_rates(&d->gates[vGate[ig]], vMin, vMax, (unsigned)vDiv, A, B);
            table_add_columns(&s->vTable, ig, A, B, gsz);
        }
        free(A); free(B);
        s->nColumn = ngate;
        s->column = (unsigned*)malloc(sizeof(unsigned) * (ngate + 1));
        for (int ig = 0; ig < ngate; ++ig) s->column[ig] = 2u * (unsigned)species[gateId[ig]];
        int maxN = 0;
        for (int ic = 0; ic < n; ++ic) if (s->caCount[ic] > maxN) maxN = s->caCount[ic];
        s->maxCaCount = maxN;
        s->caRowCompt = (or_row*)calloc(maxN + 1, sizeof(or_row));
        s->caRowIdx = (int*)malloc(sizeof(int) * (nc + 1));
        s->nZ = 0;
        for (int c = 0; c < nc; ++c)
            if (s->Zpower[c] > 0.0) s->caRowIdx[s->nZ++] = caDependIndex[c];
        free(caGate); free(vGate); free(species);
    }
    free(gateId); free(gCaDepend); free(caDependIndex);

    /* HSolveActive::readSynapses (SpikeGens) — HSolveActiveSetup.cpp:538-578 */
    s->spike = (or_spikegen*)calloc(d->n_spikegens + 1, sizeof(or_spikegen));
    s->spikeToDesc = (int*)malloc(sizeof(int) * (d->n_spikegens + 1));
    s->nSpike = 0;
    for (int ic = 0; ic < n; ++ic)
        for (int q = sgStart[s->hinesToDesc[ic]]; q < sgStart[s->hinesToDesc[ic] + 1]; ++q) {
                int k = sgList[q];
                or_spikegen* g = &s->spike[s->nSpike];
                g->compt = ic;
                g->threshold = d->spikegens[k].threshold;
                g->refractT = d->spikegens[k].refractT;
                g->edgeTriggered = d->spikegens[k].edgeTriggered;
                g->lastEvent = 0.0; g->V = 0.0; g->fired = 0;
                s->spikeToDesc[s->nSpike++] = k;
        }
    free(sgStart); free(sgList);
    /* readExternalChannels — HSolveActiveSetup.cpp:580-601 */
    s->externalCurrent = (double*)calloc(2 * (size_t)n, sizeof(double));
    s->prevExtCurr = (double*)calloc(2 * (size_t)n, sizeof(double));
    /* HHChannel2D objects and their Interpol2D tables */
    s->nTab2d = d->n_tables2d > 0 ? d->n_tables2d : 0;
    s->tab2d = (or_tab2d*)calloc(s->nTab2d + 1, sizeof(or_tab2d));
    for (int k = 0; k < s->nTab2d; ++k) {
        const hsg_table2d_desc* td = &d->tables2d[k];
        or_tab2d* t = &s->tab2d[k];
        t->nx = td->xdivs + 1;
        t->ny = td->ydivs + 1;
        t->xmin = td->xmin; t->xmax = td->xmax; t->ymin = td->ymin; t->ymax = td->ymax;
        /* Interpol2D ctor (builtins/Interpol2D.cpp:173-191) */
        t->invDx = td->xmax != td->xmin ? td->xdivs / (td->xmax - td->xmin) : 1.0;
        t->invDy = td->ymax != td->ymin ? td->ydivs / (td->ymax - td->ymin) : 1.0;
        double* copy = (double*)malloc(sizeof(double) * (size_t)t->nx * t->ny);
        memcpy(copy, td->table, sizeof(double) * (size_t)t->nx * t->ny);
        t->t = copy;
    }
    s->n2d = d->n_chans2d > 0 ? d->n_chans2d : 0;
    s->ch2d = (or_chan2d*)calloc(s->n2d + 1, sizeof(or_chan2d));
    for (int k = 0; k < s->n2d; ++k) {
        const hsg_chan2d_desc* cd = &d->chans2d[k];
        or_chan2d* c = &s->ch2d[k];
        if (cd->compartment < 0 || cd->compartment >= nd || s->descToHines[cd->compartment] < 0) {
            or_destroy(s);
            return HSG_E_MODEL;
        }
        c->compt = s->descToHines[cd->compartment];
        c->instant = cd->instant;
        const double pw[3] = {cd->Xpower, cd->Ypower, cd->Zpower};
        const double st[3] = {cd->X, cd->Y, cd->Z};
        for (int a = 0; a < 3; ++a) {
            c->pw[a] = pw[a];
            c->st[a] = st[a];
            c->tabA[a] = cd->tableA[a];
            c->tabB[a] = cd->tableB[a];
            if (pw[a] > 0.0 && (cd->tableA[a] < 0 || cd->tableA[a] >= s->nTab2d ||
                                cd->tableB[a] < 0 || cd->tableB[a] >= s->nTab2d)) {
                or_destroy(s);
                return HSG_E_MODEL;
            }
        }
        for (int q = 0; q < 6; ++q) c->dep[q] = cd->dep[q];
        c->conc1 = cd->conc1 >= 0 && cd->conc1 < d->n_pools ? s->descToPool[cd->conc1] : -1;
        c->conc2 = cd->conc2 >= 0 && cd->conc2 < d->n_pools ? s->descToPool[cd->conc2] : -1;
        if ((cd->conc1 >= 0 && c->conc1 < 0) || (cd->conc2 >= 0 && c->conc2 < 0)) {
            or_destroy(s);                /* a pool the solver does not run */
            return HSG_E_UNSUPPORTED;
        }
        c->Gbar = cd->Gbar;
        c->Ek = cd->Ek;
        c->mod = cd->modulation;
        c->Gk = 0.0;
    }
    s->valuesSent = 0;
    /* PulseGens: an injection entry on their target (HSolve::addInject
     * creates it, hsolve/HSolveInterface.cpp:213-219) */
    s->nPulse = d->n_pulsegens > 0 ? d->n_pulsegens : 0;
    s->pulse = (or_pulse*)calloc(s->nPulse + 1, sizeof(or_pulse));
    for (int k = 0; k < s->nPulse; ++k) {
        const hsg_pulsegen_desc* pd = &d->pulsegens[k];
        if ((pd->compartment >= 0 &&
             (pd->compartment >= d->n_compartments || s->descToHines[pd->compartment] < 0)) ||
            pd->compartment < -1 || pd->n_pulses < 1 || pd->n_pulses > HSG_MAX_PULSES || pd->trigMode < 0 ||
            pd->trigMode > 2 ||
            (pd->trigMode != 0 && (pd->input_src < -1 || pd->input_src >= s->nPulse || pd->input_src == k))) {
            or_destroy(s);
            return HSG_E_ARG;
        }
        or_pulse* p = &s->pulse[k];
        p->ic = pd->compartment >= 0 ? s->descToHines[pd->compartment] : -1;
        p->n = pd->n_pulses;
        p->trigMode = pd->trigMode;
        p->src = pd->trigMode != 0 ? pd->input_src : -1;
        p->base = pd->baseLevel;
        p->out = 0.0;
        p->input = 0.0;
        p->prevInput = 0.0;
        p->trigTime = -1.0;
        for (int j = 0; j < p->n; ++j) {
            p->delay[j] = pd->delay[j];
            p->width[j] = pd->width[j];
            p->level[j] = pd->level[j];
        }
        if (p->ic >= 0) s->h.hasInject[p->ic] = 1;
    }
    /* Leakage objects keep sending on their own tick */
    s->nLeak = d->n_leaks > 0 ? d->n_leaks : 0;
    s->leak = (or_leak*)calloc(s->nLeak + 1, sizeof(or_leak));
    for (int k = 0; k < s->nLeak; ++k) {
        const hsg_leak_desc* ld = &d->leaks[k];
        if (ld->compartment < 0 || ld->compartment >= d->n_compartments) {
            or_destroy(s);
            return HSG_E_ARG;
        }
        s->leak[k].ic = s->descToHines[ld->compartment];   /* -1: off the tree, not the solver's */
        s->leak[k].Gbar = ld->Gbar;
        s->leak[k].Ek = ld->Ek;
        s->leak[k].mod = ld->modulation;
    }
    /* SynChans with their handlers' synapses */
    s->nSyn = d->n_synchans > 0 ? d->n_synchans : 0;
    s->syn = (or_synchan*)calloc(s->nSyn + 1, sizeof(or_synchan));
    s->nSynapse = 0;
    for (int k = 0; k < s->nSyn; ++k) {
        const hsg_synchan_desc* sd = &d->synchans[k];
        if (sd->compartment < 0 || sd->compartment >= d->n_compartments || s->descToHines[sd->compartment] < 0 ||
            sd->n_synapses < 0) {
            or_destroy(s);
            return HSG_E_ARG;
        }
        or_synchan* c = &s->syn[k];
        c->ic = s->descToHines[sd->compartment];
        c->firstSyn = s->nSynapse;
        c->nSyn = sd->n_synapses;
        s->nSynapse += sd->n_synapses;
        c->Gbar = sd->Gbar;
        c->Ek = sd->Ek;
        c->tau1 = sd->tau1;
        c->tau2 = sd->tau2;
        c->mod = sd->modulation;
        synchan_constants(c, 25.0e-6);   /* SynChan's dt_ before any reinit */
    }
    /* VClamps: an injection entry on their compartment, as addInject
     * creates one */
    s->nVc = d->n_vclamps > 0 ? d->n_vclamps : 0;
    s->vc = (or_vclamp*)calloc(s->nVc + 1, sizeof(or_vclamp));
    for (int k = 0; k < s->nVc; ++k) {
        const hsg_vclamp_desc* vd = &d->vclamps[k];
        if (vd->compartment < 0 || vd->compartment >= d->n_compartments || s->descToHines[vd->compartment] < 0 ||
            vd->mode < 0 || vd->mode > 2) {
            or_destroy(s);
            return HSG_E_ARG;
        }
        or_vclamp* c = &s->vc[k];
        c->ic = s->descToHines[vd->compartment];
        c->mode = vd->mode;
        c->tau = vd->tau;
        c->ti = vd->ti;
        c->td = vd->td;
        c->Kp = vd->gain;
        s->h.hasInject[c->ic] = 1;
    }
    s->synW = (double*)calloc(s->nSynapse + 1, sizeof(double));
    s->synDelay = (double*)calloc(s->nSynapse + 1, sizeof(double));
    for (int j = 0; j < s->nSynapse; ++j) {
        s->synW[j] = d->synapses[j].weight;
        s->synDelay[j] = d->synapses[j].delay;
    }
    *out = s;
    return 0;
}

/* VClamp::process (biophysics/VClamp.cpp:287-311): the low-pass filtered
 * command, the error against the sensed Vm and the PID current increment */
static void vclamp_process(or_vclamp* c) {
    double dCmd = c->cmdIn - c->oldCmdIn;
    c->command = c->cmdIn + dCmd * (1 - c->tauByDt) + (c->command - c->cmdIn + dCmd * c->tauByDt) * c->expt;
    c->oldCmdIn = c->cmdIn;
    c->e = c->command - c->vIn;
    if (c->mode == 0) {
        c->current += c->Kp * ((1 + c->dtByTi + c->tdByDt) * c->e - (1 + 2 * c->tdByDt) * c->e1 + c->tdByDt * c->e2);
        c->e2 = c->e1;
        c->e1 = c->e;
    } else if (c->mode == 1) {
        c->current += c->Kp * ((1 + c->dtByTi) * c->e - c->e1 + c->tdByDt * (c->vIn - 2 * c->v1 + c->e2));
        c->e2 = c->v1;
        c->v1 = c->vIn;
        c->e1 = c->e;
    } else {
        c->current += c->Kp * (c->vIn - c->v1 + c->dtByTi * c->e + c->tdByDt * (c->vIn - 2 * c->v1 + c->e2));
        c->e2 = c->v1;
        c->v1 = c->vIn;
    }
}

/* PulseGen::process (device/PulseGen.cpp:359-424): the output at time t in
 * the PulseGen's trigger mode, from its input_; returns 0 when the reference
 * sends nothing (phase past the period), leaving out at baseLevel. */
static int pulse_process(or_pulse* p, double t) {
    double period = p->width[0] + p->delay[0];
    for (int ii = 1; ii < p->n && (p->width[ii] > 0.0 || p->delay[ii] > 0.0); ++ii) {
        double incr = p->delay[ii] + p->width[ii] - p->width[ii - 1];
        if (incr > 0) period += incr;
    }
    double phase = 0.0;
    if (p->trigMode == 0) {
        phase = fmod(t, period);
    } else if (p->trigMode == 1) {             /* EXT_TRIG */
        if (p->input == 0) {
            phase = p->trigTime < 0 ? period : t - p->trigTime;
        } else {
            if (p->prevInput == 0) p->trigTime = t;
            phase = t - p->trigTime;
        }
        p->prevInput = p->input;
    } else {                                   /* EXT_GATE */
        if (p->input == 0) {
            phase = period;
        } else {
            if (p->prevInput == 0) p->trigTime = t;
            phase = fmod(t - p->trigTime, period);
        }
        p->prevInput = p->input;
    }
    if (phase >= period) {
        p->out = p->base;
        return 0;
    }
    for (int ii = 0; ii < p->n; ++ii) {
        if (phase < p->delay[ii]) {
            p->out = p->base;
            break;
        } else if (phase < p->delay[ii] + p->width[ii]) {
            p->out = p->level[ii];
            break;
        }
        phase -= p->delay[ii];
    }
    return 1;
}

/* outputOut of PulseGen k (device/PulseGen.cpp:423, :432): to its
 * compartment's injectMsg (HSolve::addInject) and to the input of every
 * PulseGen it drives, at once, as the reference delivers messages. */
static void pulse_send(or_solver* s, int k) {
    const or_pulse* p = &s->pulse[k];
    if (p->ic >= 0) s->h.injectVarying[p->ic] += p->out;
    for (int j = 0; j < s->nPulse; ++j)
        if (s->pulse[j].src == k) s->pulse[j].input = p->out;
}

void or_destroy(or_solver* s) {
    if (!s) return;
    hm_free(&s->h);
    free(s->hinesToDesc); free(s->descToHines);
    free(s->Gbar); free(s->Xpower); free(s->Ypower); free(s->Zpower); free(s->modulation);
    free(s->instant); free(s->Gk); free(s->Ek); free(s->state); free(s->chan2state);
    free(s->chan2compt); free(s->channelCount); free(s->currentBoundary);
    free(s->chanToDesc); free(s->descToChan);
    free(s->caConc); free(s->ca); free(s->caActivation); free(s->caTarget); free(s->caCount);
    free(s->poolToDesc); free(s->descToPool);
    free(s->column); free(s->caRowIdx); free(s->caRowCompt);
    free(s->vTable.table); free(s->caTable.table);
    free(s->externalCalcium); free(s->externalCurrent); free(s->prevExtCurr);
    free(s->spike); free(s->spikeToDesc);
    iv_free(&s->spikeGenLog); free(s->spikeStepLog);
    if (s->tab2d)
        for (int k = 0; k < s->nTab2d; ++k) free((void*)s->tab2d[k].t);
    free(s->tab2d); free(s->ch2d); free(s->pulse); free(s->leak);
    for (int k = 0; k < s->nSyn; ++k) free(s->syn[k].ev);
    free(s->syn); free(s->synW); free(s->synDelay); free(s->vc);
    free(s);
}

/* ------------------------------------------------------------------------ */
/* reinit — HSolve::reinit + HSolveActive::reinit (HSolve.cpp:229-233;       */
/* HSolveActiveSetup.cpp:43-160)                                             */
/* ------------------------------------------------------------------------ */
static void reinit_channels(or_solver* s) {
    or_hines* h = &s->h;
    double* istate = s->state;
    int ichan = 0, ica = 0, icolumn = 0, icarow = 0, iextca = 0;
    or_row vRow, dRow;
    double C1, C2;
    for (unsigned iv = 0; iv < h->nCompt; ++iv) {
        table_row(&s->vTable, h->V[iv], &vRow);
        int icr = 0;
        int caBoundary = ica + s->caCount[iv];
        for (; ica < caBoundary; ++ica) { table_row(&s->caTable, s->ca[ica], &s->caRowCompt[icr]); ++icr; }
        int chanBoundary = ichan + s->channelCount[iv];
        for (; ichan < chanBoundary; ++ichan) {
            table_row(&s->caTable, s->externalCalcium[iextca], &dRow);
            if (s->Xpower[ichan] > 0.0) {
                table_lookup(&s->vTable, s->column[icolumn], &vRow, &C1, &C2);
                *istate = C1 / C2;
                ++icolumn, ++istate;
            }
            if (s->Ypower[ichan] > 0.0) {
                table_lookup(&s->vTable, s->column[icolumn], &vRow, &C1, &C2);
                *istate = C1 / C2;
                ++icolumn, ++istate;
            }
            if (s->Zpower[ichan] > 0.0) {
                int cr = s->caRowIdx[icarow];
                if (cr >= 0) table_lookup(&s->caTable, s->column[icolumn], &s->caRowCompt[cr], &C1, &C2);
                else if (s->externalCalcium[iextca] > 0) table_lookup(&s->caTable, s->column[icolumn], &dRow, &C1, &C2);
                else table_lookup(&s->vTable, s->column[icolumn], &vRow, &C1, &C2);
                *istate = C1 / C2;
                ++icolumn, ++istate, ++icarow;
            }
            ++iextca;
        }
    }
}

static void chan2d_reinit(or_solver* s);

void or_reinit(or_solver* s, double dt) {
    or_hines* h = &s->h;
    /* PulseGen::reinit (device/PulseGen.cpp:426-433) on tick 0 sends
     * baseLevel; the solver's reinit leaves injectVarying alone, so it is
     * added to the first step */
    for (int k = 0; k < s->nPulse; ++k) {
        or_pulse* p = &s->pulse[k];
        p->trigTime = -1.0;
        p->prevInput = 0.0;
        p->out = p->base;
        p->input = 0.0;
        pulse_send(s, k);
    }
    /* VClamp::reinit (VClamp.cpp:313-347, tick 0): the error terms and
     * the command filter restart at the compartment's initVm; the current
     * is kept */
    for (int k = 0; k < s->nVc; ++k) {
        or_vclamp* c = &s->vc[k];
        c->vIn = 0.0;
        c->v1 = 0.0;
        c->command = c->cmdIn = c->oldCmdIn = c->e = c->e1 = c->e2 = 0.0;
        if (c->ti == 0.0) c->ti = dt;
        if (c->td < 0.0) c->td = 0.0;
        if (c->tau == 0.0) c->tau = 5.0 * dt;
        c->expt = (dt / c->tau > 1e-15) ? exp(-dt / c->tau) : 1.0 - dt / c->tau;
        c->tauByDt = c->tau / dt;
        c->dtByTi = dt / c->ti;
        c->tdByDt = c->td / dt;
        if (c->Kp == 0.0) c->Kp = h->Cm[c->ic] / dt;
        c->command = c->cmdIn = c->oldCmdIn = h->initVm[c->ic];
    }
    chan2d_reinit(s);            /* their reinit runs before the solver's */
    /* SimpleSynHandler::vReinit drops the queue; SynChan::vReinit
     * (SynChan.cpp:229-251) zeroes X, Y, Gk, activation at the new dt */
    for (int k = 0; k < s->nSyn; ++k) {
        or_synchan* c = &s->syn[k];
        c->nEv = 0;
        c->activation = 0.0;
        c->X = c->Y = c->Gk = 0.0;
        synchan_constants(c, dt);
    }
    s->hsolveDt = dt;
    for (unsigned i = 0; i < 2 * h->nCompt; ++i) s->externalCurrent[i] = 0.0;
    for (int k = 0; k < s->nSpike; ++k) s->spike[k].lastEvent = -s->spike[k].refractT;
    for (unsigned ic = 0; ic < h->nCompt; ++ic) h->V[ic] = h->initVm[ic];
    for (int p = 0; p < s->nPool; ++p) {
        s->caActivation[p] = 0.0;
        s->caConc[p].c = 0.0;
        s->ca[p] = s->caConc[p].CaBasal;
    }
    reinit_channels(s);
    s->stepCount = 0;
    s->spikeGenLog.n = 0;
    s->valuesSent = 1;           /* HSolveActive::reinit ends with sendValues */
}

/* ------------------------------------------------------------------------ */
/* step — HSolveActive::step (HSolveActive.cpp:45-66)                        */
/* ------------------------------------------------------------------------ */

/* advanceChannels — HSolveActive.cpp:218-332 */
static void advance_channels(or_solver* s, double dt) {
    or_hines* h = &s->h;
    double* istate = s->state;
    int ichan = 0, ica = 0, icolumn = 0, icarow = 0, iextca = 0;
    or_row vRow, dRow;
    double C1 = 0.0, C2 = 0.0;
    for (unsigned iv = 0; iv < h->nCompt; ++iv) {
        table_row(&s->vTable, h->V[iv], &vRow);
        int icr = 0;
        int caBoundary = ica + s->caCount[iv];
        for (; ica < caBoundary; ++ica) { table_row(&s->caTable, s->ca[ica], &s->caRowCompt[icr]); ++icr; }
        int chanBoundary = ichan + s->channelCount[iv];
        for (; ichan < chanBoundary; ++ichan) {
            table_row(&s->caTable, s->externalCalcium[iextca], &dRow);
            if (s->Xpower[ichan] > 0.0) {
                table_lookup(&s->vTable, s->column[icolumn], &vRow, &C1, &C2);
                if (s->instant[ichan] & 1) *istate = C1 / C2;
                else {
                    double temp = 1.0 + dt / 2.0 * C2;
                    *istate = (*istate * (2.0 - temp) + dt * C1) / temp;
                }
                ++icolumn, ++istate;
            }
            if (s->Ypower[ichan] > 0.0) {
                table_lookup(&s->vTable, s->column[icolumn], &vRow, &C1, &C2);
                if (s->instant[ichan] & 2) *istate = C1 / C2;
                else {
                    double temp = 1.0 + dt / 2.0 * C2;
                    *istate = (*istate * (2.0 - temp) + dt * C1) / temp;
                }
                ++icolumn, ++istate;
            }
            if (s->Zpower[ichan] > 0.0) {
                int cr = s->caRowIdx[icarow];
                if (cr >= 0) table_lookup(&s->caTable, s->column[icolumn], &s->caRowCompt[cr], &C1, &C2);
                else if (s->externalCalcium[iextca] > 0) table_lookup(&s->caTable, s->column[icolumn], &dRow, &C1, &C2);
                else table_lookup(&s->vTable, s->column[icolumn], &vRow, &C1, &C2);
                if (s->instant[ichan] & 4) *istate = C1 / C2;
                else {
                    double temp = 1.0 + dt / 2.0 * C2;
                    *istate = (*istate * (2.0 - temp) + dt * C1) / temp;
                }
                ++icolumn, ++istate, ++icarow;
            }
            ++iextca;
        }
    }
}

/* ChannelStruct::process + powers — HSolveStruct.cpp:15-65 */
static double take_power(double x, double p) {
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p == 3.0) return x * x * x;
    if (p == 4.0) { double y = x * x; return y * y; }
    if (x > 0.0) return exp(p * log(x));
    return 0.0;
}

/* calculateChannelCurrents — HSolveActive.cpp:68-83 */
static void calculate_channel_currents(or_solver* s) {
    const double* istate = s->state;
    if (s->nState == 0) return;
    for (int c = 0; c < s->nChan; ++c) {
        double fraction = s->modulation[c];
        if (s->Xpower[c] > 0.0) fraction *= take_power(*(istate++), s->Xpower[c]);
        if (s->Ypower[c] > 0.0) fraction *= take_power(*(istate++), s->Ypower[c]);
        if (s->Zpower[c] > 0.0) fraction *= take_power(*(istate++), s->Zpower[c]);
        s->Gk[c] = s->Gbar[c] * fraction;
    }
}

/* HSolveActive::updateMatrix — HSolveActive.cpp:85-151 */
static void update_matrix_active(or_solver* s) {
    or_hines* h = &s->h;
    if (h->nHJ) memcpy(h->HJ, h->HJCopy, sizeof(double) * h->nHJ);
    int icur = 0;
    for (unsigned ic = 0; ic < h->nCompt; ++ic) {
        double GkSum = 0.0, GkEkSum = 0.0;
        for (; icur < s->currentBoundary[ic]; ++icur) {
            GkSum += s->Gk[icur];
            GkEkSum += s->Gk[icur] * s->Ek[icur];
        }
        double* ihs = h->HS + 4 * ic;
        ihs[0] = ihs[2] + GkSum;
        ihs[3] = h->V[ic] * h->CmByDt[ic] + h->EmByRm[ic] + GkEkSum;
    }
    for (unsigned ic = 0; ic < h->nCompt; ++ic) {
        if (!h->hasInject[ic]) continue;
        h->HS[4 * ic + 3] += h->injectVarying[ic] + h->injectBasal[ic];
        h->injectVarying[ic] = 0.0;
    }
    for (unsigned ic = 0; ic < h->nCompt; ++ic) {
        h->HS[4 * ic] += s->externalCurrent[2 * ic];
        h->HS[4 * ic + 3] += s->externalCurrent[2 * ic + 1];
    }
    h->stage = 0;
}

/* CaConcStruct::process — HSolveStruct.cpp:138-155 */
static double caconc_process(or_caconc* c, double activation) {
    c->c = c->factor1 * c->c + c->factor2 * activation;
    double ca = c->CaBasal + c->c;
    if (c->ceiling > 0 && ca > c->ceiling) { ca = c->ceiling; c->c = ca - c->CaBasal; }
    if (ca < c->floor) { ca = c->floor; c->c = ca - c->CaBasal; }
    return ca;
}

/* advanceCalcium — HSolveActive.cpp:153-216 */
static void advance_calcium(or_solver* s) {
    or_hines* h = &s->h;
    int icur = 0, ict = 0;
    if (s->caAdvance == 1) {
        for (unsigned ib = 0; ib < h->nCompt; ++ib) {
            for (; icur < s->currentBoundary[ib]; ++icur) {
                if (s->caTarget[ict] >= 0)
                    s->caActivation[s->caTarget[ict]] += s->Gk[icur] * (s->Ek[icur] - h->VMid[ib]);
                ++ict;
            }
        }
    } else if (s->caAdvance == 0) {
        for (unsigned ib = 0; ib < h->nCompt; ++ib) {
            for (; icur < s->currentBoundary[ib]; ++icur) {
                if (s->caTarget[ict] >= 0) {
                    double v0 = (2 * h->VMid[ib] - h->V[ib]);
                    s->caActivation[s->caTarget[ict]] += s->Gk[icur] * (s->Ek[icur] - v0);
                }
                ++ict;
            }
        }
    }
    for (int p = 0; p < s->nPool; ++p) s->ca[p] = caconc_process(&s->caConc[p], s->caActivation[p]);
    for (int p = 0; p < s->nPool; ++p) s->caActivation[p] = 0.0;
}

/* SpikeGen::process via SpikeGenStruct::send — SpikeGen.cpp:202-216 */
static void send_spikes(or_solver* s, double dt, double t) {
    for (int k = 0; k < s->nSpike; ++k) {
        or_spikegen* g = &s->spike[k];
        g->V = s->h.V[g->compt];
        if (g->V > g->threshold) {
            if ((t + dt / 2.0) >= (g->lastEvent + g->refractT)) {
                if (!(g->edgeTriggered && g->fired)) {
                    iv_push(&s->spikeGenLog, k);
                    if (s->spikeGenLog.n > s->spikeStepCap) {
                        s->spikeStepCap = s->spikeGenLog.cap;
                        s->spikeStepLog = (long*)realloc(s->spikeStepLog, sizeof(long) * s->spikeStepCap);
                    }
                    s->spikeStepLog[s->spikeGenLog.n - 1] = s->stepCount;
                    g->lastEvent = t;
                    g->fired = 1;
                }
            }
        } else {
            g->fired = 0;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* HHChannel2D beside the solver                                             */
/* ------------------------------------------------------------------------ */

/* Interpol2D::innerLookup and ::interpolate (builtins/Interpol2D.cpp:508-590) */
static double interp2d(const or_tab2d* t, double x, double y) {
    if (t->nx == 0) return 0.0;
    if (x < t->xmin) x = t->xmin;
    if (x > t->xmax) x = t->xmax;
    if (y < t->ymin) y = t->ymin;
    if (y > t->ymax) y = t->ymax;
    int endX = 0, endY = 0;
    double z00 = 0.0, z01 = 0.0, z10 = 0.0, z11 = 0.0;
    double xv = (x - t->xmin) * t->invDx;
    unsigned long xi = (unsigned long)xv;
    if (xi >= t->nx) xi = t->nx - 1;
    if (xi == t->nx - 1) endX = 1;
    double xF = xv - xi;
    double yv = (y - t->ymin) * t->invDy;
    unsigned long yi = (unsigned long)yv;
    if (yi >= t->ny) yi = t->ny - 1;
    if (yi == t->ny - 1) endY = 1;
    double yF = yv - yi;
    double xFyF = xF * yF;
    z00 = t->t[xi * t->ny + yi];
    if (!endX) {
        z10 = t->t[(xi + 1) * t->ny + yi];
        if (!endY) {
            z11 = t->t[(xi + 1) * t->ny + yi + 1];
            z01 = t->t[xi * t->ny + yi + 1];
        }
    } else if (!endY) {
        z01 = t->t[xi * t->ny + yi + 1];
    }
    return z00 * (1 - xF - yF + xFyF) + z10 * (xF - xFyF) + z01 * (yF - xFyF) + z11 * xFyF;
}

/* HHChannelBase::integrate (biophysics/HHChannelBase.cpp:268-274) */
static double integrate_ee(double state, double dt, double A, double B) {
    if (B > 1.0e-10) {
        double x = exp(-B * dt);
        return state * x + (A / B) * (1 - x);
    }
    return state + A * dt;
}

/* HHChannel2D::depValue (HHChannel2D.cpp:257-271): the Vm and concentrations
 * the channel last received (zero before the solver first sends them) */
static double dep_value(const or_solver* s, const or_chan2d* c, int dep) {
    if (!s->valuesSent) return 0.0;
    switch (dep) {
    case 0: return s->h.V[c->compt];
    case 1: return c->conc1 >= 0 ? s->ca[c->conc1] : 0.0;
    case 2: return c->conc2 >= 0 ? s->ca[c->conc2] : 0.0;
    default: return 0.0;
    }
}

/* HHChannel2D::vProcess (HHChannel2D.cpp:325-368), then the channel's
 * channelOut -> HSolve::addGkEk (HSolveInterface.cpp:246-252) */
static void chan2d_process(or_solver* s, double dt) {
    for (int k = 0; k < s->n2d; ++k) {
        or_chan2d* c = &s->ch2d[k];
        double g = 0.0;
        g += c->Gbar;
        for (int a = 0; a < 3; ++a) {
            if (!(c->pw[a] > 0.0)) continue;
            const double u = dep_value(s, c, c->dep[2 * a]), w = dep_value(s, c, c->dep[2 * a + 1]);
            const double A = interp2d(&s->tab2d[c->tabA[a]], u, w);
            const double B = interp2d(&s->tab2d[c->tabB[a]], u, w);
            if (c->instant & (1 << a)) c->st[a] = A / B;
            else c->st[a] = integrate_ee(c->st[a], dt, A, B);
            g *= take_power(c->st[a], c->pw[a]);
        }
        c->Gk = g * c->mod;
        s->externalCurrent[2 * c->compt] += c->Gk;
        s->externalCurrent[2 * c->compt + 1] += c->Gk * c->Ek;
    }
}

/* HHChannel2D::vReinit (HHChannel2D.cpp:374-426): gates at A/B of the last
 * received values; a table B below EPSILON stops the channel's reinit */
static void chan2d_reinit(or_solver* s) {
    for (int k = 0; k < s->n2d; ++k) {
        or_chan2d* c = &s->ch2d[k];
        double g = c->Gbar;
        int stop = 0;
        for (int a = 0; a < 3 && !stop; ++a) {
            if (!(c->pw[a] > 0.0)) continue;
            const double u = dep_value(s, c, c->dep[2 * a]), w = dep_value(s, c, c->dep[2 * a + 1]);
            const double A = interp2d(&s->tab2d[c->tabA[a]], u, w);
            const double B = interp2d(&s->tab2d[c->tabB[a]], u, w);
            if (B < 1.0e-10) { stop = 1; break; }
            c->st[a] = A / B;
            g *= take_power(c->st[a], c->pw[a]);
        }
        if (!stop) c->Gk = g * c->mod;
    }
}

void or_step(or_solver* s, double dt, double currTime) {
    or_hines* h = &s->h;
    if (h->nCompt <= 0) return;
    s->stepCount++;
    /* PulseGens (tick 0), then the HHChannel2D objects (tick 4), before the
     * solver (tick 6) */
    for (int k = 0; k < s->nPulse; ++k)
        if (pulse_process(&s->pulse[k], currTime)) pulse_send(s, k);
    /* VClamps (tick 0, after the PulseGens), from the Vm the solver sent at
     * the end of the previous tick */
    for (int k = 0; k < s->nVc; ++k) {
        or_vclamp* c = &s->vc[k];
        c->vIn = h->V[c->ic];
        vclamp_process(c);
        h->injectVarying[c->ic] += c->current;
    }
    /* SimpleSynHandler::vProcess (tick 1, SimpleSynHandler.cpp:104-112) */
    for (int k = 0; k < s->nSyn; ++k) {
        or_synchan* c = &s->syn[k];
        double activation = 0.0;
        int q = 0;
        while (q < c->nEv && c->ev[q].time <= currTime) {
            activation += c->ev[q].weight / dt;
            ++q;
        }
        if (q > 0) {
            memmove(c->ev, c->ev + q, sizeof(or_event) * (size_t)(c->nEv - q));
            c->nEv -= q;
        }
        if (activation != 0.0) c->activation += activation;
    }
    /* tick 4: Leakage (Leakage.cpp:90-95), the HHChannel2D objects, the
     * SynChans (SynChan.cpp:200-227), each sending (Gk, Ek) to addGkEk */
    for (int k = 0; k < s->nLeak; ++k) {
        const or_leak* l = &s->leak[k];
        if (l->ic < 0) continue;
        const double Gk = l->Gbar * l->mod;
        s->externalCurrent[2 * l->ic] += Gk;
        s->externalCurrent[2 * l->ic + 1] += Gk * l->Ek;
    }
    chan2d_process(s, dt);
    for (int k = 0; k < s->nSyn; ++k) {
        or_synchan* c = &s->syn[k];
        c->X = c->activation * c->xconst1 + c->X * c->xconst2;
        c->Y = c->X * c->yconst1 + c->Y * c->yconst2;
        c->activation = 0.0;
        c->Gk = c->Y * c->norm * c->mod;
        s->externalCurrent[2 * c->ic] += c->Gk;
        s->externalCurrent[2 * c->ic + 1] += c->Gk * c->Ek;
    }
    advance_channels(s, dt);
    calculate_channel_currents(s);
    update_matrix_active(s);
    hp_forward(h);
    hp_backward(h);
    advance_calcium(s);
    send_spikes(s, dt, currTime);
    memcpy(s->prevExtCurr, s->externalCurrent, sizeof(double) * 2 * h->nCompt);
    for (unsigned i = 0; i < 2 * h->nCompt; ++i) s->externalCurrent[i] = 0.0;
    s->valuesSent = 1;           /* sendValues */
}

void or_run(or_solver* s, double dt, long nsteps, const int* probe, int nprobe, double* trace) {
    for (long k = 0; k < nsteps; ++k) {
        /* Clock::handleStep: currTime = dt * endStep (scheduling/Clock.cpp:726-727) */
        double t = dt * (double)(unsigned long)(s->stepCount + 1);
        or_step(s, dt, t);
        if (trace)
            for (int j = 0; j < nprobe; ++j) trace[k * nprobe + j] = s->h.V[s->descToHines[probe[j]]];
    }
}

/* ------------------------------------------------------------------------ */
/* accessors — HSolveInterface.cpp                                           */
/* ------------------------------------------------------------------------ */
int or_n_compartments(const or_solver* s) { return (int)s->h.nCompt; }
int or_n_channels(const or_solver* s) { return s->nChan; }
int or_n_states(const or_solver* s) { return s->nState; }
int or_n_pools(const or_solver* s) { return s->nPool; }
void or_hines_order(const or_solver* s, int* out) {
    for (unsigned i = 0; i < s->h.nCompt; ++i) out[s->hinesToDesc[i]] = (int)i;
}
void or_get_V(const or_solver* s, double* out) { memcpy(out, s->h.V, sizeof(double) * s->h.nCompt); }
void or_get_VMid(const or_solver* s, double* out) { memcpy(out, s->h.VMid, sizeof(double) * s->h.nCompt); }
void or_get_state(const or_solver* s, double* out) { memcpy(out, s->state, sizeof(double) * s->nState); }
void or_get_ca(const or_solver* s, double* out) { memcpy(out, s->ca, sizeof(double) * s->nPool); }

static int state_index(const or_solver* s, int c, int which) {
    /* getX/getY/getZ — HSolveInterface.cpp:341-443 */
    int si = s->chan2state[c];
    if (which >= 1 && s->Xpower[c] > 0.0) ++si;
    if (which >= 2 && s->Ypower[c] > 0.0) ++si;
    return si;
}

int or_get_field(const or_solver* s, int field, const int* idx, int n, double* out) {
    const or_hines* h = &s->h;
    for (int k = 0; k < n; ++k) {
        int i = idx[k];
        if (field < 20) {
            if (i < 0 || i >= (int)h->nCompt || s->descToHines[i] < 0) return HSG_E_ARG;
            int ic = s->descToHines[i];
            switch (field) {
            case HSG_F_VM: out[k] = h->V[ic]; break;
            case HSG_F_IM: {
                double Im = h->EmByRm[ic] - h->V[ic] / h->Rm[ic];
                int start = ic == 0 ? 0 : s->currentBoundary[ic - 1];
                for (int c = start; c < s->currentBoundary[ic]; ++c) Im += (s->Ek[c] - h->V[ic]) * s->Gk[c];
                Im += s->prevExtCurr[2 * ic + 1] - s->prevExtCurr[2 * ic] * h->V[ic];
                out[k] = Im;
                break;
            }
            case HSG_F_INJECT: out[k] = h->hasInject[ic] ? h->injectBasal[ic] : 0.0; break;
            case HSG_F_INITVM: out[k] = h->initVm[ic]; break;
            case HSG_F_CM: out[k] = h->Cm[ic]; break;
            case HSG_F_EM: out[k] = h->Em[ic]; break;
            case HSG_F_RM: out[k] = h->Rm[ic]; break;
            case HSG_F_RA: out[k] = h->Ra[ic]; break;
            default: return HSG_E_ARG;
            }
        } else if (field < 40) {
            if (i < 0 || s->descToChan[i] < 0) return HSG_E_ARG;
            int c = s->descToChan[i];
            switch (field) {
            case HSG_F_GBAR: out[k] = s->Gbar[c]; break;
            case HSG_F_EK: out[k] = s->Ek[c]; break;
            case HSG_F_GK: out[k] = s->Gk[c]; break;
            case HSG_F_IK: out[k] = (s->Ek[c] - h->V[s->chan2compt[c]]) * s->Gk[c]; break;
            case HSG_F_X: out[k] = s->Xpower[c] == 0.0 ? 0.0 : s->state[state_index(s, c, 0)]; break;
            case HSG_F_Y: out[k] = s->Ypower[c] == 0.0 ? 0.0 : s->state[state_index(s, c, 1)]; break;
            case HSG_F_Z: out[k] = s->Zpower[c] == 0.0 ? 0.0 : s->state[state_index(s, c, 2)]; break;
            case HSG_F_MODULATION: out[k] = s->modulation[c]; break;
            case HSG_F_INSTANT: out[k] = (double)s->instant[c]; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 110) {
            if (i < 0 || i >= s->nPulse) return HSG_E_ARG;
            switch (field) {
            case HSG_F_PULSE_OUTPUT: out[k] = s->pulse[i].out; break;
            case HSG_F_PULSE_INPUT: out[k] = s->pulse[i].input; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 100) {
            if (i < 0 || i >= s->nVc) return HSG_E_ARG;
            switch (field) {
            case HSG_F_VCLAMP_COMMAND: out[k] = s->vc[i].cmdIn; break;
            case HSG_F_VCLAMP_CURRENT: out[k] = s->vc[i].current; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 95) {
            if (i < 0 || i >= s->nLeak) return HSG_E_ARG;
            const or_leak* l = &s->leak[i];
            switch (field) {
            case HSG_F_LEAK_GBAR: out[k] = l->Gbar; break;
            case HSG_F_LEAK_EK: out[k] = l->Ek; break;
            case HSG_F_LEAK_MODULATION: out[k] = l->mod; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 90) {
            if (i < 0 || i >= s->nSynapse) return HSG_E_ARG;
            switch (field) {
            case HSG_F_SYNAPSE_WEIGHT: out[k] = s->synW[i]; break;
            case HSG_F_SYNAPSE_DELAY: out[k] = s->synDelay[i]; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 80) {
            if (i < 0 || i >= s->nSyn) return HSG_E_ARG;
            const or_synchan* c = &s->syn[i];
            switch (field) {
            case HSG_F_SYN_GBAR: out[k] = c->Gbar; break;
            case HSG_F_SYN_EK: out[k] = c->Ek; break;
            case HSG_F_SYN_GK: out[k] = c->Gk; break;
            case HSG_F_SYN_TAU1: out[k] = c->tau1; break;
            case HSG_F_SYN_TAU2: out[k] = c->tau2; break;
            case HSG_F_SYN_MODULATION: out[k] = c->mod; break;
            case HSG_F_SYN_X: out[k] = c->X; break;
            case HSG_F_SYN_Y: out[k] = c->Y; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 60) {
            if (i < 0 || i >= s->n2d) return HSG_E_ARG;
            const or_chan2d* c = &s->ch2d[i];
            switch (field) {
            case HSG_F_2D_GBAR: out[k] = c->Gbar; break;
            case HSG_F_2D_EK: out[k] = c->Ek; break;
            case HSG_F_2D_GK: out[k] = c->Gk; break;
            case HSG_F_2D_X: out[k] = c->st[0]; break;
            case HSG_F_2D_Y: out[k] = c->st[1]; break;
            case HSG_F_2D_Z: out[k] = c->st[2]; break;
            case HSG_F_2D_MODULATION: out[k] = c->mod; break;
            default: return HSG_E_ARG;
            }
        } else {
            if (i < 0 || s->descToPool[i] < 0) return HSG_E_ARG;
            int p = s->descToPool[i];
            switch (field) {
            case HSG_F_CA: out[k] = s->ca[p]; break;
            case HSG_F_CABASAL: out[k] = s->caConc[p].CaBasal; break;
            case HSG_F_CACEIL: out[k] = s->caConc[p].ceiling; break;
            case HSG_F_CAFLOOR: out[k] = s->caConc[p].floor; break;
            default: return HSG_E_ARG;
            }
        }
    }
    return 0;
}

int or_set_field(or_solver* s, int field, const int* idx, int n, const double* v) {
    or_hines* h = &s->h;
    for (int k = 0; k < n; ++k) {
        int i = idx[k];
        double x = v[k];
        if (field < 20) {
            if (i < 0 || s->descToHines[i] < 0) return HSG_E_ARG;
            int ic = s->descToHines[i];
            switch (field) {
            case HSG_F_VM: h->V[ic] = x; break;
            case HSG_F_INJECT: h->hasInject[ic] = 1; h->injectBasal[ic] = x; break;
            case HSG_F_INITVM: h->initVm[ic] = x; break;
            case HSG_F_CM: h->Cm[ic] = x; h->CmByDt[ic] = 2.0 * x / s->hsolveDt; break;
            case HSG_F_EM: h->Em[ic] = x; h->EmByRm[ic] = x / h->Rm[ic]; break;
            case HSG_F_RM: h->Rm[ic] = x; h->EmByRm[ic] = h->Em[ic] / x; break;
            case HSG_F_RA: h->Ra[ic] = x; break;
            default: return HSG_E_ARG;
            }
        } else if (field < 40) {
            if (i < 0 || s->descToChan[i] < 0) return HSG_E_ARG;
            int c = s->descToChan[i];
            switch (field) {
            case HSG_F_GBAR: s->Gbar[c] = x; break;
            case HSG_F_EK: s->Ek[c] = x; break;
            case HSG_F_GK: s->Gk[c] = x; break;
            case HSG_F_X: if (s->Xpower[c] != 0.0) s->state[state_index(s, c, 0)] = x; break;
            case HSG_F_Y: if (s->Ypower[c] != 0.0) s->state[state_index(s, c, 1)] = x; break;
            case HSG_F_Z: if (s->Zpower[c] != 0.0) s->state[state_index(s, c, 2)] = x; break;
            case HSG_F_MODULATION: if (x > 0.0) s->modulation[c] = x; break;
            case HSG_F_INSTANT: s->instant[c] = (int)x; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 110) {
            /* PulseGen::input (device/PulseGen.cpp:355-358) where no
             * PulseGen's output drives it */
            if (i < 0 || i >= s->nPulse || field != HSG_F_PULSE_INPUT || s->pulse[i].src >= 0) return HSG_E_ARG;
            s->pulse[i].input = x;
        } else if (field >= 100) {
            if (i < 0 || i >= s->nVc) return HSG_E_ARG;
            switch (field) {
            case HSG_F_VCLAMP_COMMAND: s->vc[i].cmdIn = x; break;   /* VClamp::setCommand */
            default: return HSG_E_ARG;
            }
        } else if (field >= 95) {
            if (i < 0 || i >= s->nLeak) return HSG_E_ARG;
            or_leak* l = &s->leak[i];
            switch (field) {
            case HSG_F_LEAK_GBAR: l->Gbar = x; break;
            case HSG_F_LEAK_EK: l->Ek = x; break;
            case HSG_F_LEAK_MODULATION: l->mod = x; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 90) {
            if (i < 0 || i >= s->nSynapse) return HSG_E_ARG;
            switch (field) {
            case HSG_F_SYNAPSE_WEIGHT: s->synW[i] = x; break;
            case HSG_F_SYNAPSE_DELAY: s->synDelay[i] = x; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 80) {
            if (i < 0 || i >= s->nSyn) return HSG_E_ARG;
            or_synchan* c = &s->syn[i];
            switch (field) {
            /* the setters recompute the constants at the object's dt_
             * (SynChan.cpp:117-156) */
            case HSG_F_SYN_GBAR: c->Gbar = x; synchan_constants(c, c->dt); break;
            case HSG_F_SYN_TAU1: c->tau1 = x; synchan_constants(c, c->dt); break;
            case HSG_F_SYN_TAU2: c->tau2 = x; synchan_constants(c, c->dt); break;
            case HSG_F_SYN_EK: c->Ek = x; break;
            case HSG_F_SYN_MODULATION: c->mod = x; break;
            default: return HSG_E_ARG;
            }
        } else if (field >= 60) {
            if (i < 0 || i >= s->n2d) return HSG_E_ARG;
            or_chan2d* c = &s->ch2d[i];
            switch (field) {
            case HSG_F_2D_GBAR: c->Gbar = x; break;
            case HSG_F_2D_EK: c->Ek = x; break;
            case HSG_F_2D_X: c->st[0] = x; break;
            case HSG_F_2D_Y: c->st[1] = x; break;
            case HSG_F_2D_Z: c->st[2] = x; break;
            case HSG_F_2D_MODULATION: c->mod = x; break;
            default: return HSG_E_ARG;
            }
        } else {
            if (i < 0 || s->descToPool[i] < 0) return HSG_E_ARG;
            int p = s->descToPool[i];
            switch (field) {
            case HSG_F_CA:
            case HSG_F_CABASAL:   /* ZombieCaConc::vSetCaBasal calls setCa (ZombieCaConc.cpp:62-65) */
                s->ca[p] = x;
                s->caConc[p].c = x - s->caConc[p].CaBasal;
                break;
            case HSG_F_CACEIL: s->caConc[p].ceiling = x; break;
            case HSG_F_CAFLOOR: s->caConc[p].floor = x; break;
            default: return HSG_E_ARG;
            }
        }
    }
    return 0;
}

int or_add_inject(or_solver* s, int compartment, double value) {
    if (compartment < 0 || s->descToHines[compartment] < 0) return HSG_E_ARG;
    int ic = s->descToHines[compartment];
    s->h.hasInject[ic] = 1;
    s->h.injectVarying[ic] += value;
    return 0;
}
/* Synapse::addSpike (synapse/Synapse.cpp:104-113) into the handler's
 * priority queue (SimpleSynHandler::addSpike, :93-97) */
int or_add_spike(or_solver* s, int synchan, int synapse, double time) {
    if (synchan < 0 || synchan >= s->nSyn) return HSG_E_ARG;
    or_synchan* c = &s->syn[synchan];
    if (synapse < 0 || synapse >= c->nSyn) return HSG_E_ARG;
    const int j = c->firstSyn + synapse;
    or_event e = {time + s->synDelay[j], s->synW[j]};
    if (c->nEv == c->capEv) {
        c->capEv = c->capEv ? 2 * c->capEv : 16;
        c->ev = (or_event*)realloc(c->ev, sizeof(or_event) * (size_t)c->capEv);
    }
    int q = c->nEv;
    while (q > 0 && c->ev[q - 1].time > e.time) {
        c->ev[q] = c->ev[q - 1];
        --q;
    }
    c->ev[q] = e;
    c->nEv++;
    return 0;
}

int or_add_gkek(or_solver* s, int compartment, double Gk, double Ek) {
    if (compartment < 0 || s->descToHines[compartment] < 0) return HSG_E_ARG;
    int ic = s->descToHines[compartment];
    s->externalCurrent[2 * ic] += Gk;
    s->externalCurrent[2 * ic + 1] += Gk * Ek;
    return 0;
}
int or_add_conc(or_solver* s, int channel, double conc) {
    if (channel < 0 || s->descToChan[channel] < 0) return HSG_E_ARG;
    s->externalCalcium[s->descToChan[channel]] = conc;
    return 0;
}
int or_add_ica(or_solver* s, int pool, double iCa) {
    if (pool < 0 || s->descToPool[pool] < 0) return HSG_E_ARG;
    s->caActivation[s->descToPool[pool]] += iCa;
    return 0;
}
int or_read_spikes(or_solver* s, int* gen, long* step, int cap) {
    int n = s->spikeGenLog.n < cap ? s->spikeGenLog.n : cap;
    for (int k = 0; k < n; ++k) { gen[k] = s->spikeToDesc[s->spikeGenLog.a[k]]; step[k] = s->spikeStepLog[k]; }
    s->spikeGenLog.n = 0;
    return n;
}

/* ------------------------------------------------------------------------ */
/* flat dump for the reference harness (oracle/ref_harness.cpp reads it)     */
/* ------------------------------------------------------------------------ */
static void wr_i(FILE* f, long x) { fwrite(&x, sizeof(long), 1, f); }
static void wr_d(FILE* f, double x) { fwrite(&x, sizeof(double), 1, f); }
static void wr_dv(FILE* f, const double* a, long n) { wr_i(f, n); if (n) fwrite(a, sizeof(double), (size_t)n, f); }
static void wr_iv(FILE* f, const int* a, long n) { wr_i(f, n); for (long k = 0; k < n; ++k) wr_i(f, a[k]); }

int or_dump_flat(const or_solver* s, const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) return HSG_E_ARG;
    const or_hines* h = &s->h;
    long n = h->nCompt;
    wr_i(f, 0x48534f4cL);          /* magic */
    wr_i(f, n);
    wr_d(f, h->dt);
    wr_i(f, s->caAdvance);
    wr_iv(f, h->cstart, n + 1);
    wr_iv(f, h->children, h->cstart[n]);
    wr_dv(f, h->Ra, n); wr_dv(f, h->Rm, n); wr_dv(f, h->Cm, n); wr_dv(f, h->Em, n); wr_dv(f, h->initVm, n);
    wr_dv(f, h->V, n); wr_dv(f, h->CmByDt, n); wr_dv(f, h->EmByRm, n);
    wr_iv(f, h->hasInject, n); wr_dv(f, h->injectBasal, n); wr_dv(f, h->injectVarying, n);
    long nc = s->nChan;
    wr_i(f, nc);
    wr_dv(f, s->Gbar, nc); wr_dv(f, s->Xpower, nc); wr_dv(f, s->Ypower, nc); wr_dv(f, s->Zpower, nc);
    wr_iv(f, s->instant, nc); wr_dv(f, s->modulation, nc); wr_dv(f, s->Ek, nc);
    wr_dv(f, s->state, s->nState);
    wr_iv(f, s->chan2state, nc); wr_iv(f, s->chan2compt, nc); wr_iv(f, s->channelCount, n);
    long np = s->nPool;
    wr_i(f, np);
    for (long p = 0; p < np; ++p) {
        wr_d(f, s->caConc[p].c); wr_d(f, s->caConc[p].CaBasal); wr_d(f, s->caConc[p].factor1);
        wr_d(f, s->caConc[p].factor2); wr_d(f, s->caConc[p].ceiling); wr_d(f, s->caConc[p].floor);
    }
    wr_dv(f, s->ca, np);
    wr_iv(f, s->caTarget, nc);
    wr_iv(f, s->caCount, n);
    {
        int* col = (int*)malloc(sizeof(int) * (s->nColumn + 1));
        for (int k = 0; k < s->nColumn; ++k) col[k] = (int)s->column[k];
        wr_iv(f, col, s->nColumn);
        free(col);
    }
    wr_iv(f, s->caRowIdx, s->nZ);
    wr_dv(f, s->externalCalcium, nc);
    const or_table* tabs[2] = {&s->vTable, &s->caTable};
    for (int t = 0; t < 2; ++t) {
        const or_table* tb = tabs[t];
        wr_d(f, tb->min); wr_d(f, tb->max);
        wr_i(f, (long)(tb->nPts - 2));
        wr_i(f, (long)(tb->nColumns / 2));
        wr_dv(f, tb->table, (long)tb->nPts * tb->nColumns);
    }
    fclose(f);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* bare HinesMatrix for the matrix-level tests                               */
/* ------------------------------------------------------------------------ */
struct or_matrix { or_hines h; };

or_matrix* or_matrix_create(int n, const int* cstart, const int* children,
                            const double* Ra, const double* Rm, const double* Cm,
                            const double* Em, const double* V, double dt) {
    or_matrix* m = (or_matrix*)calloc(1, sizeof(or_matrix));
    or_hines* h = &m->h;
    h->nCompt = (unsigned)n;
    h->dt = dt;
    h->cstart = (int*)malloc(sizeof(int) * (n + 1));
    memcpy(h->cstart, cstart, sizeof(int) * (n + 1));
    h->children = (int*)malloc(sizeof(int) * (cstart[n] + 1));
    memcpy(h->children, children, sizeof(int) * cstart[n]);
    h->Ra = (double*)malloc(sizeof(double) * n); memcpy(h->Ra, Ra, sizeof(double) * n);
    h->Rm = (double*)malloc(sizeof(double) * n); memcpy(h->Rm, Rm, sizeof(double) * n);
    h->Cm = (double*)malloc(sizeof(double) * n); memcpy(h->Cm, Cm, sizeof(double) * n);
    h->Em = (double*)malloc(sizeof(double) * n); memcpy(h->Em, Em, sizeof(double) * n);
    h->initVm = (double*)malloc(sizeof(double) * n); memcpy(h->initVm, V, sizeof(double) * n);
    h->V = (double*)malloc(sizeof(double) * n); memcpy(h->V, V, sizeof(double) * n);
    h->CmByDt = (double*)malloc(sizeof(double) * n);
    h->EmByRm = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) { h->CmByDt[i] = 2.0 * Cm[i] / dt; h->EmByRm[i] = Em[i] / Rm[i]; }
    h->hasInject = (int*)calloc(n, sizeof(int));
    h->injectVarying = (double*)calloc(n, sizeof(double));
    h->injectBasal = (double*)calloc(n, sizeof(double));
    hm_setup(h);
    return m;
}
void or_matrix_destroy(or_matrix* m) { if (m) { hm_free(&m->h); free(m); } }
double or_matrix_getA(const or_matrix* m, int row, int col) { return hm_getA(&m->h, (unsigned)row, (unsigned)col); }
double or_matrix_getB(const or_matrix* m, int row) { return m->h.HS[4 * row + 3]; }
double or_matrix_getVMid(const or_matrix* m, int row) { return m->h.VMid[row]; }
double or_matrix_getV(const or_matrix* m, int row) { return m->h.V[row]; }
void or_matrix_update(or_matrix* m) { hp_update_matrix(&m->h); }
void or_matrix_forward(or_matrix* m) { hp_forward(&m->h); }
void or_matrix_backward(or_matrix* m) { hp_backward(&m->h); }

/* Single-lookup entry points of the 2-D table and the exponential-Euler
 * step, for the closed-form checks of the tests. */
double or_interp2d(const hsg_table2d_desc* td, double x, double y) {
    or_tab2d t;
    t.nx = td->xdivs + 1;
    t.ny = td->ydivs + 1;
    t.xmin = td->xmin; t.xmax = td->xmax; t.ymin = td->ymin; t.ymax = td->ymax;
    t.invDx = td->xmax != td->xmin ? td->xdivs / (td->xmax - td->xmin) : 1.0;
    t.invDy = td->ymax != td->ymin ? td->ydivs / (td->ymax - td->ymin) : 1.0;
    t.t = (double*)td->table;
    return interp2d(&t, x, y);
}

double or_integrate_ee(double state, double dt, double A, double B) { return integrate_ee(state, dt, A, B); }
