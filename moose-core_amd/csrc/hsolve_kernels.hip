// hsolve_kernels.hip — the per-step HSolve path on CDNA4 (gfx950).
//
// One workgroup advances one cell for `nsteps` consecutive ticks
// (HSolveActive::step, hsolve/HSolveActive.cpp:45-66, repeated).  Cells are
// independent, so a launch needs no grid-wide synchronisation between steps.
// Per tick the workgroup runs:
//   A  channels + assembly: one thread per compartment (strided), gate update
//      from the lookup tables (advanceChannels, HSolveActive.cpp:218-332),
//      Gk (calculateChannelCurrents, :68-83) and the matrix row
//      (updateMatrix, :85-151) written to the elimination work arrays.
//   F  forward elimination over the unit schedule, one unit per thread and a
//      barrier per level (forwardEliminate, hsolve/HSolvePassive.cpp:246-323).
//   B  back-substitution in reverse level order (backwardSubstitute,
//      HSolvePassive.cpp:325-401) and V = 2*VMid - V.
//   D  Ca pools (advanceCalcium, HSolveActive.cpp:153-216), spike generators
//      (SpikeGen::process, biophysics/SpikeGen.cpp:202-216) and Vm probes.
// A "unit" is either a linear link i -> i+1 or a junction group (all members
// but the last eliminated, in member order, into the later members).  Units of
// one level touch disjoint rows.  Terms that the reference applies to a row
// only after that row's earlier group siblings are parked in a deferred-term
// slot and applied in the reference's order, so that the FP64 form (built with
// -ffp-contract=off and IEEE division) reproduces the reference bit for bit.
//
// The solve is latency-bound, so: the elimination work arrays live in LDS,
// the barriers between levels order LDS only (the next level's unit
// descriptors, loaded from HBM, stay in flight across them), the cell's
// fields and table parameters sit in registers, and the channel types of a
// uniform cell sit in LDS.
//
// GT is the channel type: double (FP64 form) or float (mixed form, where the
// rate tables and their interpolation, Gbar, Ek, the gate powers and the
// channel currents are FP32).  Gate states and their trapezoidal update, the
// Hines solve and Vm are FP64 in both forms: a gate state stored in FP32
// loses the per-step increments that are below half its ulp (near rest they
// are), and the bias this leaves moves spike times by several steps.
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#include "hsolve_internal.h"

namespace hsg {

template <typename T>
__device__ __forceinline__ T* A_(char* arena, int64_t off) {
    return reinterpret_cast<T*>(arena + off);
}

// Barrier that orders LDS only: global loads issued before it stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS-only barrier when the work arrays are in LDS, a full one otherwise.
__device__ __forceinline__ void work_barrier(bool ldsWork) {
    if (ldsWork) lds_barrier();
    else __syncthreads();
}

// One lookup table in registers (hsolve/RateLookup.h LookupTable).
template <typename GT>
struct Table {
    double min, max, dx, rdx;   // rdx = 1 / dx
    int nCol;
    const GT* __restrict__ tab;
};

template <typename GT>
__device__ __forceinline__ Table<GT> table_view(char* arena, const TableDev& t) {
    Table<GT> v;
    v.min = t.min;
    v.max = t.max;
    v.dx = t.dx;
    v.rdx = 1.0 / t.dx;
    v.nCol = t.nColumns;
    v.tab = reinterpret_cast<const GT*>(arena + (sizeof(GT) == 8 ? t.off64 : t.off32));
    return v;
}

// LookupTable::row (hsolve/RateLookup.cpp:63-76).  The mixed form (GT =
// float) multiplies by 1 / dx instead of dividing: the row and fraction then
// differ from the quotient's by at most an ulp of the position, which moves
// the interpolated value by less than the FP32 table's own rounding.
template <typename GT>
__device__ __forceinline__ void table_row(const Table<GT>& t, double x, int& base, double& frac) {
    if (x < t.min) x = t.min;
    else if (x > t.max) x = t.max;
    double div = std::is_same<GT, float>::value ? (x - t.min) * t.rdx : (x - t.min) / t.dx;
    unsigned integer = (unsigned)div;
    frac = div - integer;
    base = (int)integer * t.nCol;
}

// LookupTable::lookup (hsolve/RateLookup.cpp:78-103)
template <typename GT>
__device__ __forceinline__ void table_lookup(const Table<GT>& t, int base, int col, GT frac, GT& C1, GT& C2) {
    const GT* ap = t.tab + base + col;
    const GT* bp = ap + t.nCol;
    GT a = ap[0], b = bp[0];
    C1 = a + (b - a) * frac;
    a = ap[1];
    b = bp[1];
    C2 = a + (b - a) * frac;
}

// ChannelStruct powers (hsolve/HSolveStruct.cpp:15-65, HSolveStruct.h:88-108).
// Powers 1-4 by multiplication; any other power through exp(p*log(x)), kept
// out of line so that its registers do not count against the hot path.
__device__ __noinline__ double power_general(double x, double p) {
    if (x > 0.0) return exp(p * log(x));
    return 0.0;
}

template <typename GT>
__device__ __forceinline__ GT take_power(GT x, double p) {
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p == 3.0) return x * x * x;
    if (p == 4.0) {
        GT y = x * x;
        return y * y;
    }
    return (GT)power_general((double)x, p);
}

// Quotient a / b.  FAST = false: the IEEE quotient (the FP64 form reproduces
// the reference bit for bit).  FAST = true (the mixed form): the hardware
// reciprocal (relative error up to 2^-24 on gfx950, measured over 1e-14..1e14)
// refined by one Newton step (below 2^-48 measured) times a: within about
// ten ulps of the IEEE quotient for operands of ordinary range, in four
// dependent steps instead of about ten.
template <bool FAST>
__device__ __forceinline__ double quot(double a, double b) {
    if constexpr (!FAST) {
        return a / b;
    } else {
        double r = __builtin_amdgcn_rcp(b);
        const double e = fma(-b, r, 1.0);
        r = fma(e, r, r);
        return a * r;
    }
}

// trapezoidal gate update, dx/dt = A - B x (HSolveActive.cpp:259-268), in
// FP64; FAST (the mixed form) divides with quot<true>
template <bool FAST>
__device__ __forceinline__ double gate_step(double x, double C1, double C2, double dt, bool instant) {
    if (instant) return quot<FAST>(C1, C2);
    double temp = 1.0 + dt / 2.0 * C2;
    return quot<FAST>(x * (2.0 - temp) + dt * C1, temp);
}

struct Work {
    double* D;
    double* B;     // becomes VMid during back-substitution
    double* HJ;    // working junction coefficients of slow-path groups (HBM, or LDS in blocks)
    const double* HJ0;   // their initial values (HJCopy_): HBM, or the same LDS array in blocks
    double* T;     // deferred terms: pairs (tD, tB)
    double* F;     // fill entry A[1][2] of each three-member unit
};

// Everything a workgroup reads about its cell, hoisted into registers.
template <typename GT>
struct View {
    int n, nChanSlots, uniform, hasModulation, hasPools, caAdvance;
    double* __restrict__ V;
    const double* __restrict__ CmByDt;
    const double* __restrict__ EmByRm;
    const double* __restrict__ Dbase;
    double* __restrict__ state;   // gate states: FP64 in both forms
    const GT* __restrict__ Gbar;
    const GT* __restrict__ Ek;
    const double* __restrict__ mod;
    const double* __restrict__ extCa;
    const int32_t* __restrict__ chType;
    const int32_t* __restrict__ poolCount;
    double* __restrict__ poolCa;
    Table<GT> vt, ct;
};

template <typename GT>
__device__ __forceinline__ View<GT> make_view(char* arena, const CellDev& c, const TableSetDev* tsets) {
    View<GT> v;
    v.n = c.n;
    v.nChanSlots = c.nChanSlots;
    v.uniform = c.uniform;
    v.hasModulation = c.hasModulation;
    v.hasPools = c.hasPools;
    v.caAdvance = c.caAdvance;
    v.V = A_<double>(arena, c.oV);
    v.CmByDt = A_<const double>(arena, c.oCmByDt);
    v.EmByRm = A_<const double>(arena, c.oEmByRm);
    v.Dbase = A_<const double>(arena, c.oDbase);
    v.state = A_<double>(arena, c.oState);
    v.Gbar = A_<const GT>(arena, c.oGbar);
    v.Ek = A_<const GT>(arena, c.oEk);
    v.mod = A_<const double>(arena, c.oMod);
    v.extCa = A_<const double>(arena, c.oExtCa);
    v.chType = A_<const int32_t>(arena, c.oChType);
    v.poolCount = A_<const int32_t>(arena, c.oPoolCount);
    v.poolCa = A_<double>(arena, c.oPoolCa);
    const TableSetDev& ts = tsets[c.tableSet];
    v.vt = table_view<GT>(arena, ts.v);
    v.ct = table_view<GT>(arena, ts.ca);
    return v;
}

// ---------------------------------------------------------------------------
// phase A: channels and matrix row of compartment i.  `types` is the cell's
// slot-type table in LDS (uniform cells) or the engine's type table.
// ---------------------------------------------------------------------------
template <typename GT, bool UNIFORM>
__device__ __forceinline__ void channel_row(const View<GT>& v, const ChanType* __restrict__ types, int i,
                                            double dtg, Work w, bool reinitOnly, int dst) {
    const int n = v.n;
    const double V = v.V[i];
    // pools of this compartment; the Ca row of a pool (caRowCompt_) is
    // computed where a Z gate reads it
    const int nPools = v.hasPools ? v.poolCount[i] : 0;
    int vBase;
    double vFrac;
    table_row(v.vt, V, vBase, vFrac);
    const GT frac = (GT)vFrac;
    double GkSum = 0.0, GkEkSum = 0.0;
    int js = 0;   // gate-state slot
    for (int k = 0; k < v.nChanSlots; ++k) {
        const int t = UNIFORM ? k : v.chType[(size_t)k * n + i];
        if (t < 0) continue;
        const ChanType& ct = types[t];
        if (UNIFORM && ct.nGates < 0) continue;   // empty slot
        GT C1, C2;
        GT fraction = v.hasModulation ? (GT)v.mod[(size_t)k * n + i] : (GT)1.0;
        if (ct.Xpower > 0.0) {
            double* sp = v.state + (size_t)js * n + i;
            table_lookup<GT>(v.vt, vBase, ct.colX, frac, C1, C2);
            const double x = reinitOnly ? (double)C1 / (double)C2
                                        : gate_step<std::is_same<GT, float>::value>(*sp, C1, C2, dtg, ct.instant & 1);
            *sp = x;
            fraction *= take_power<GT>((GT)x, ct.Xpower);
            ++js;
        }
        if (ct.Ypower > 0.0) {
            double* sp = v.state + (size_t)js * n + i;
            table_lookup<GT>(v.vt, vBase, ct.colY, frac, C1, C2);
            const double x = reinitOnly ? (double)C1 / (double)C2
                                        : gate_step<std::is_same<GT, float>::value>(*sp, C1, C2, dtg, ct.instant & 2);
            *sp = x;
            fraction *= take_power<GT>((GT)x, ct.Ypower);
            ++js;
        }
        if (ct.Zpower > 0.0) {
            double* sp = v.state + (size_t)js * n + i;
            if (ct.caDepLocal >= 0 && ct.caDepLocal < nPools) {
                int cb;
                double cf;
                table_row(v.ct, v.poolCa[(size_t)ct.caDepLocal * n + i], cb, cf);
                table_lookup<GT>(v.ct, cb, ct.colZ, (GT)cf, C1, C2);
            } else {
                const double ext = v.extCa[(size_t)k * n + i];
                if (ext > 0 && ct.colZ + 1 < v.ct.nCol) {
                    int b;
                    double f;
                    table_row(v.ct, ext, b, f);
                    table_lookup<GT>(v.ct, b, ct.colZ, (GT)f, C1, C2);
                } else {
                    table_lookup<GT>(v.vt, vBase, ct.colZ, frac, C1, C2);
                }
            }
            const double x = reinitOnly ? (double)C1 / (double)C2
                                        : gate_step<std::is_same<GT, float>::value>(*sp, C1, C2, dtg, ct.instant & 4);
            *sp = x;
            fraction *= take_power<GT>((GT)x, ct.Zpower);
            ++js;
        }
        if (!reinitOnly) {
            GT Gk = v.Gbar[(size_t)k * n + i] * fraction;
            GkSum += (double)Gk;
            GkEkSum += (double)Gk * (double)v.Ek[(size_t)k * n + i];
        }
    }
    if (reinitOnly) return;
    // updateMatrix row (HSolveActive.cpp:112-115)
    w.D[dst] = v.Dbase[i] + GkSum;
    w.B[dst] = V * v.CmByDt[i] + v.EmByRm[i] + GkEkSum;
}

// ---------------------------------------------------------------------------
// HHChannel2D objects beside the solver (biophysics/HHChannel2D.cpp).  They
// tick before the solver from the Vm and [Ca] it sent at the end of the
// previous tick (zero before it first sends them), integrate their gates by
// exponential Euler and add (Gk, Gk*Ek) to their compartment's external
// current, which updateMatrix adds to the row in one sum
// (HSolveActive.cpp:140-148).  Computed in FP64 in both forms.
// ---------------------------------------------------------------------------

// Interpol2D::innerLookup and ::interpolate (builtins/Interpol2D.cpp:508-590)
__device__ __forceinline__ double interp2d(const Table2DDev& t, const double* __restrict__ tab, double x, double y) {
    if (x < t.xmin) x = t.xmin;
    if (x > t.xmax) x = t.xmax;
    if (y < t.ymin) y = t.ymin;
    if (y > t.ymax) y = t.ymax;
    const unsigned long nx = (unsigned long)t.nx, ny = (unsigned long)t.ny;
    const double xv = (x - t.xmin) * t.invDx;
    unsigned long xi = (unsigned long)xv;
    if (xi >= nx) xi = nx - 1;
    const bool endX = xi == nx - 1;
    const double xF = xv - xi;
    const double yv = (y - t.ymin) * t.invDy;
    unsigned long yi = (unsigned long)yv;
    if (yi >= ny) yi = ny - 1;
    const bool endY = yi == ny - 1;
    const double yF = yv - yi;
    const double xFyF = xF * yF;
    const double z00 = tab[xi * ny + yi];
    double z01 = 0.0, z10 = 0.0, z11 = 0.0;
    if (!endX) {
        z10 = tab[(xi + 1) * ny + yi];
        if (!endY) {
            z11 = tab[(xi + 1) * ny + yi + 1];
            z01 = tab[xi * ny + yi + 1];
        }
    } else if (!endY) {
        z01 = tab[xi * ny + yi + 1];
    }
    return z00 * (1 - xF - yF + xFyF) + z10 * (xF - xFyF) + z01 * (yF - xFyF) + z11 * xFyF;
}

// Both tables of a gate at once from its pair record (the A and B tables
// share one grid, so Interpol2D's cell and weights are shared): the same
// index arithmetic as interp2d, corner pairs and the weighted sums in GT
// (FP64: the same operations as two interp2d calls; mixed form: FP32 values
// and weights).
template <typename GT>
__device__ __forceinline__ void interp2d_pair(const Table2DDev& t, const char* arena, double x, double y, double& A,
                                              double& B) {
    using V2 = typename std::conditional<std::is_same<GT, float>::value, float2, double2>::type;
    const V2* __restrict__ tab = A_<const V2>(const_cast<char*>(arena), std::is_same<GT, float>::value ? t.off32 : t.off);
    if (x < t.xmin) x = t.xmin;
    if (x > t.xmax) x = t.xmax;
    if (y < t.ymin) y = t.ymin;
    if (y > t.ymax) y = t.ymax;
    const unsigned long nx = (unsigned long)t.nx, ny = (unsigned long)t.ny;
    const double xv = (x - t.xmin) * t.invDx;
    unsigned long xi = (unsigned long)xv;
    if (xi >= nx) xi = nx - 1;
    const bool endX = xi == nx - 1;
    const double xFd = xv - xi;
    const double yv = (y - t.ymin) * t.invDy;
    unsigned long yi = (unsigned long)yv;
    if (yi >= ny) yi = ny - 1;
    const bool endY = yi == ny - 1;
    const double yFd = yv - yi;
    const GT xF = (GT)xFd, yF = (GT)yFd;
    const GT xFyF = xF * yF;
    const V2 z00 = tab[xi * ny + yi];
    V2 z01, z10, z11;
    z01.x = z01.y = z10.x = z10.y = z11.x = z11.y = (GT)0;
    if (!endX) {
        z10 = tab[(xi + 1) * ny + yi];
        if (!endY) {
            z11 = tab[(xi + 1) * ny + yi + 1];
            z01 = tab[xi * ny + yi + 1];
        }
    } else if (!endY) {
        z01 = tab[xi * ny + yi + 1];
    }
    const GT w00 = 1 - xF - yF + xFyF, w10 = xF - xFyF, w01 = yF - xFyF;
    A = (double)(z00.x * w00 + z10.x * w10 + z01.x * w01 + z11.x * xFyF);
    B = (double)(z00.y * w00 + z10.y * w10 + z01.y * w01 + z11.y * xFyF);
}

// HHChannelBase::integrate (biophysics/HHChannelBase.cpp:268-274)
__device__ __forceinline__ double integrate_ee(double state, double dt, double A, double B) {
    if (B > 1.0e-10) {
        const double x = exp(-B * dt);
        return state * x + (A / B) * (1 - x);
    }
    return state + A * dt;
}

// HHChannel2D::depValue (HHChannel2D.cpp:257-271)
__device__ __forceinline__ double dep2d(int dep, bool sent, double Vm, const Chan2DDev& ch,
                                        const double* __restrict__ poolCa) {
    if (!sent) return 0.0;
    if (dep == 0) return Vm;
    if (dep == 1) return ch.conc1 >= 0 ? poolCa[ch.conc1] : 0.0;
    if (dep == 2) return ch.conc2 >= 0 ? poolCa[ch.conc2] : 0.0;
    return 0.0;
}

// The objects that send (Gk, Ek) to compartment i on every tick before the
// solver's tick (clock tick 4), summed into the compartment's external
// current (HSolve::addGkEk, hsolve/HSolveInterface.cpp:246-252) after the
// call's host entries (first tick of a call): the Leakage objects
// (Leakage::vProcess, biophysics/Leakage.cpp:90-95), the HHChannel2D
// objects (HHChannel2D::vProcess, :325-368) and the SynChans
// (SynChan::vProcess, biophysics/SynChan.cpp:200-227, with the activation
// their handlers sent on tick 1), each kind in description order.  The sum
// is kept in o2dExt for the row and for Im.  Run by its own kernel before
// the tick's solver kernels, so the step kernels carry none of its
// registers.  F32 (mixed form): the HHChannel2D tables are read as FP32
// pairs and Gk is formed in FP32; the gate state and its integration stay
// FP64.
template <bool WITH2D, bool F32 = false>
__device__ __forceinline__ double2 external_sum(char* arena, const Table2DDev* __restrict__ tabs2d, double dt,
                                                bool first, const CellDev& c, int i, double Vm,
                                                const double* __restrict__ poolCa, int64_t stepNo) {
    double G = 0.0, GE = 0.0;
    if (first) {
        const int32_t* xi = A_<const int32_t>(arena, c.oExtIdx);
        for (int e = 0; e < c.nExt; ++e)
            if (xi[e] == i) {
                G += A_<const double>(arena, c.oExtG)[e];
                GE += A_<const double>(arena, c.oExtGE)[e];
            }
    }
    if (c.nLeak > 0) {
        const int32_t* ls = A_<const int32_t>(arena, c.oLeakStart);
        const LeakDev* lk = A_<const LeakDev>(arena, c.oLeak);
#pragma unroll 1
        for (int k = ls[i]; k < ls[i + 1]; ++k) {
            G += lk[k].G;
            GE += lk[k].GE;
        }
    }
    if (WITH2D && c.n2d > 0) {
        const int32_t* start = A_<const int32_t>(arena, c.o2dStart);
        const Chan2DDev* chans = A_<const Chan2DDev>(arena, c.o2dChan);
        const Chan2DType* kinds = A_<const Chan2DType>(arena, c.o2dType);
        double* st = A_<double>(arena, c.o2dState);
        double* gk = A_<double>(arena, c.o2dGk);
        const bool sent = c.values2dSent != 0;
#pragma unroll 1
        for (int k = start[i]; k < start[i + 1]; ++k) {
            const Chan2DDev& ch = chans[k];
            const Chan2DType& ty = kinds[ch.type];
            using CT = typename std::conditional<F32, float, double>::type;
            CT g = 0.0;
            g += (CT)ch.Gbar;
#pragma unroll 1
            for (int ax = 0; ax < 3; ++ax) {
                if (!(ty.pw[ax] > 0.0)) continue;
                const double u = dep2d(ty.dep[2 * ax], sent, Vm, ch, poolCa);
                const double w = dep2d(ty.dep[2 * ax + 1], sent, Vm, ch, poolCa);
                double A, B;
                if (ty.pair[ax] >= 0) {
                    interp2d_pair<CT>(tabs2d[ty.pair[ax]], arena, u, w, A, B);
                } else {
                    const Table2DDev& ta = tabs2d[ty.tabA[ax]];
                    const Table2DDev& tb = tabs2d[ty.tabB[ax]];
                    A = interp2d(ta, A_<const double>(arena, ta.off), u, w);
                    B = interp2d(tb, A_<const double>(arena, tb.off), u, w);
                }
                double x = st[3 * k + ax];
                x = (ty.instant >> ax) & 1 ? A / B : integrate_ee(x, dt, A, B);
                st[3 * k + ax] = x;
                g *= take_power<CT>((CT)x, ty.pw[ax]);
            }
            const double Gk = (double)(g * (CT)ch.mod);
            gk[k] = Gk;
            G += Gk;
            GE += Gk * ch.Ek;
        }
    }
    if (c.nSyn > 0) {
        const int32_t* ss = A_<const int32_t>(arena, c.oSynStart);
        SynDev* sy = A_<SynDev>(arena, c.oSyn);
        int32_t* cur = A_<int32_t>(arena, c.oActCur);
        const int32_t* end = A_<const int32_t>(arena, c.oActEnd);
        const ActDev* act = A_<const ActDev>(arena, c.oAct);
#pragma unroll 1
        for (int k = ss[i]; k < ss[i + 1]; ++k) {
            SynDev& q = sy[k];
            double a = 0.0;
            const int j = cur[k];
            if (j < end[k] && act[j].step == stepNo) {
                a = act[j].value;
                cur[k] = j + 1;
            }
            const double X = a * q.xconst1 + q.X * q.xconst2;
            const double Y = X * q.yconst1 + q.Y * q.yconst2;
            const double Gk = Y * q.norm * q.mod;
            q.X = X;
            q.Y = Y;
            q.Gk = Gk;
            G += Gk;
            GE += Gk * q.Ek;
        }
    }
    double* ext = A_<double>(arena, c.o2dExt);
    ext[2 * i] = G;
    ext[2 * i + 1] = GE;
    return make_double2(G, GE);
}

// external_sum for a cell whose objects beside the solver are one
// HHChannel2D per compartment, instance i on compartment i, all of one kind
// (c.uni2d): the same arithmetic in the same order, with the kind read once
// (every lane reads the same record) and the instance, its gate states and
// the compartment's own pool concentration loaded together, so that the
// table corners are the only loads that wait on others.
template <bool F32>
__device__ __forceinline__ double2 external_uni2d(char* arena, const Table2DDev* __restrict__ tabs2d, double dt,
                                                  bool first, const CellDev& c, int i, double Vm,
                                                  const double* __restrict__ poolCa) {
    using CT = typename std::conditional<F32, float, double>::type;
    const Chan2DType& ty = A_<const Chan2DType>(arena, c.o2dType)[0];
    const Chan2DDev ch = A_<const Chan2DDev>(arena, c.o2dChan)[i];
    double* st = A_<double>(arena, c.o2dState);
    double x3[3] = {st[3 * i], st[3 * i + 1], st[3 * i + 2]};
    // conc1 / conc2 are the compartment's own pool in the usual case
    const double own = c.nPoolSlots > 0 ? poolCa[i] : 0.0;
    const bool sent = c.values2dSent != 0;
    auto dep = [&](int d) -> double {
        if (!sent) return 0.0;
        if (d == 0) return Vm;
        const int o = d == 1 ? ch.conc1 : d == 2 ? ch.conc2 : -1;
        if (o < 0) return 0.0;
        return o == i ? own : poolCa[o];
    };
    double G = 0.0, GE = 0.0;
    if (first) {
        const int32_t* xi = A_<const int32_t>(arena, c.oExtIdx);
        for (int e = 0; e < c.nExt; ++e)
            if (xi[e] == i) {
                G += A_<const double>(arena, c.oExtG)[e];
                GE += A_<const double>(arena, c.oExtGE)[e];
            }
    }
    CT g = 0.0;
    g += (CT)ch.Gbar;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        if (!(ty.pw[ax] > 0.0)) continue;
        const double u = dep(ty.dep[2 * ax]);
        const double w = dep(ty.dep[2 * ax + 1]);
        double A, B;
        if (ty.pair[ax] >= 0) {
            interp2d_pair<CT>(tabs2d[ty.pair[ax]], arena, u, w, A, B);
        } else {
            const Table2DDev& ta = tabs2d[ty.tabA[ax]];
            const Table2DDev& tb = tabs2d[ty.tabB[ax]];
            A = interp2d(ta, A_<const double>(arena, ta.off), u, w);
            B = interp2d(tb, A_<const double>(arena, tb.off), u, w);
        }
        double x = x3[ax];
        x = (ty.instant >> ax) & 1 ? A / B : integrate_ee(x, dt, A, B);
        st[3 * i + ax] = x;
        g *= take_power<CT>((CT)x, ty.pw[ax]);
    }
    const double Gk = (double)(g * (CT)ch.mod);
    A_<double>(arena, c.o2dGk)[i] = Gk;
    G += Gk;
    GE += Gk * ch.Ek;
    double* ext = A_<double>(arena, c.o2dExt);
    ext[2 * i] = G;
    ext[2 * i + 1] = GE;
    return make_double2(G, GE);
}

// The external kernel's form (every kind of object).
__device__ __forceinline__ void external_objects(char* arena, const Table2DDev* __restrict__ tabs2d, double dt,
                                                 bool first, bool f32, const CellDev& c, int i, double Vm,
                                                 const double* __restrict__ poolCa, int64_t stepNo) {
    if (c.uni2d) {
        if (f32) external_uni2d<true>(arena, tabs2d, dt, first, c, i, Vm, poolCa);
        else external_uni2d<false>(arena, tabs2d, dt, first, c, i, Vm, poolCa);
        return;
    }
    if (!A_<const int32_t>(arena, c.oXFlag)[i]) return;
    if (f32) external_sum<true, true>(arena, tabs2d, dt, first, c, i, Vm, poolCa, stepNo);
    else external_sum<true, false>(arena, tabs2d, dt, first, c, i, Vm, poolCa, stepNo);
}

// The step kernels' form, for cells without HHChannel2D objects (Leakage
// and SynChan only): called where the tick's row takes its external
// current, so these cells keep multi-tick launches.  Not inlined, so the
// step kernels' register budget does not grow.
__device__ __noinline__ double2 external_light(char* arena, bool first, const CellDev& c, int i, int64_t stepNo) {
    return external_sum<false>(arena, nullptr, 0.0, first, c, i, 0.0, nullptr, stepNo);
}

// The external current of compartment i (which has objects sending every
// tick) for the tick ending at step stepNo: computed here, or read from
// o2dExt where the external kernel ran before the tick (HHChannel2D cells).
__device__ __forceinline__ double2 external_current(char* arena, bool first, const CellDev& c, int i,
                                                    int64_t stepNo) {
    if (c.n2d > 0) {
        const double* ext = A_<const double>(arena, c.o2dExt);
        return make_double2(ext[2 * i], ext[2 * i + 1]);
    }
    return external_light(arena, first, c, i, stepNo);
}

// HHChannel2D::vReinit (HHChannel2D.cpp:374-426), run before the solver's
// reinit: gates at A/B of the values last received; a table B below
// EPSILON stops that channel's reinit.
__device__ __forceinline__ void chan2d_reinit(char* arena, const Table2DDev* __restrict__ tabs2d, const CellDev& c) {
    const Chan2DDev* chans = A_<const Chan2DDev>(arena, c.o2dChan);
    const Chan2DType* kinds = A_<const Chan2DType>(arena, c.o2dType);
    double* st = A_<double>(arena, c.o2dState);
    double* gk = A_<double>(arena, c.o2dGk);
    const double* V = A_<const double>(arena, c.oV);
    const double* poolCa = A_<const double>(arena, c.oPoolCa);
    const bool sent = c.values2dSent != 0;
    for (int k = threadIdx.x; k < c.n2d; k += blockDim.x) {
        const Chan2DDev& ch = chans[k];
        const Chan2DType& ty = kinds[ch.type];
        double g = ch.Gbar;
        bool stop = false;
#pragma unroll 1
        for (int ax = 0; ax < 3 && !stop; ++ax) {
            if (!(ty.pw[ax] > 0.0)) continue;
            const double u = dep2d(ty.dep[2 * ax], sent, V[ch.compt], ch, poolCa);
            const double w = dep2d(ty.dep[2 * ax + 1], sent, V[ch.compt], ch, poolCa);
            const Table2DDev& ta = tabs2d[ty.tabA[ax]];
            const Table2DDev& tb = tabs2d[ty.tabB[ax]];
            const double A = interp2d(ta, A_<const double>(arena, ta.off), u, w);
            const double B = interp2d(tb, A_<const double>(arena, tb.off), u, w);
            if (B < 1.0e-10) {
                stop = true;
                break;
            }
            st[3 * k + ax] = A / B;
            g *= take_power<double>(st[3 * k + ax], ty.pw[ax]);
        }
        if (!stop) gk[k] = g * ch.mod;
    }
}

// ---------------------------------------------------------------------------
// elimination units
// ---------------------------------------------------------------------------
struct Desc {
    int4 a, b, c;
};

__device__ __forceinline__ Desc load_desc(const int4* __restrict__ p, int u) {
    return Desc{p[3 * u], p[3 * u + 1], p[3 * u + 2]};
}

__device__ __forceinline__ int dword(const Desc& u, int i) {
    const int4& v = i < 4 ? u.a : i < 8 ? u.b : u.c;
    const int j = i & 3;
    return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
}

__device__ __forceinline__ int member(const Desc& u, int p) {
    const int w = dword(u, 3 + p / 2);
    return (p & 1) ? ((w >> 16) & 0xffff) : (w & 0xffff);
}

__device__ __forceinline__ double desc_double(const Desc& u, int i) {
    return __hiloint2double(dword(u, i + 1), dword(u, i));
}

__device__ __forceinline__ double desc_off(const Desc& u) { return desc_double(u, 4); }

// Units of two or three members (every linear link, and the junction groups
// of a binary tree) on one code path, all arithmetic in registers.  Member 0
// is eliminated into the later members; in a three-member unit member 1 is
// then pivoted and eliminated into member 2 (the operand sequences of
// HinesMatrix::makeOperands, hsolve/HinesMatrix.cpp:225-365, and of
// HSolvePassive::forwardEliminate, hsolve/HSolvePassive.cpp:262-298, ranks 1
// and 2 and the linear run).
template <bool FD = false>
__device__ __forceinline__ void fwd_small(const Desc& u, Work w) {
    const bool three = (u.a.x & 15) == 3;
    const int m0 = u.a.w & 0xffff, m1 = (u.a.w >> 16) & 0xffff, m2 = u.b.x;
    const int deferLast = u.a.y;
    const int cnt = three ? (u.a.x >> 8) & 15 : 0;
    const int tb = u.a.z;
    const double a01 = desc_double(u, 6);
    // Every row and term the unit reads, read before anything is written, so
    // that one LDS round trip brings them all: no other unit of the level
    // writes them (a row has one feeder and one consumer, at different
    // levels), and the terms of member 1's feeder (at most two for a binary
    // junction; more in the loop below) were written at earlier levels.
    const double D0 = w.D[m0], B0 = w.B[m0];
    double D1 = w.D[m1], B1 = w.B[m1];
    double D2 = 0.0, B2 = 0.0, t0D = 0.0, t0B = 0.0, t1D = 0.0, t1B = 0.0;
    if (three && deferLast < 0) {
        D2 = w.D[m2];
        B2 = w.B[m2];
    }
    if (cnt > 0) {
        t0D = w.T[2 * tb];
        t0B = w.T[2 * tb + 1];
    }
    if (cnt > 1) {
        t1D = w.T[2 * tb + 2];
        t1B = w.T[2 * tb + 3];
    }
    const double l1 = quot<FD>(a01, D0);
    const double tD1 = l1 * a01;
    const double tB1 = l1 * B0;
    // terms for the last member, in the reference's order
    double lD1 = tD1, lB1 = tB1, lD2 = 0.0, lB2 = 0.0;
    if (three) {
        const double a02 = desc_double(u, 8), a12 = desc_double(u, 10);
        D1 -= tD1;
        B1 -= tB1;
        // member 1's feeder terms come after member 0's (deferred slot)
        if (cnt > 0) {
            D1 -= t0D;
            B1 -= t0B;
        }
        if (cnt > 1) {
            D1 -= t1D;
            B1 -= t1B;
        }
        for (int t = 2; t < cnt; ++t) {
            D1 -= w.T[2 * (tb + t)];
            B1 -= w.T[2 * (tb + t) + 1];
        }
        const double l2 = quot<FD>(a02, D0);
        const double tD2 = l2 * a02;
        const double tB2 = l2 * B0;
        const double a12u = a12 - l1 * a02;   // A[1][2] after member 0
        const double a12l = a12 - l2 * a01;   // A[2][1] after member 0
        const double l = quot<FD>(a12l, D1);
        w.D[m1] = D1;
        w.B[m1] = B1;
        w.F[u.b.y] = a12u;
        lD1 = tD2;
        lB1 = tB2;
        lD2 = l * a12u;
        lB2 = l * B1;
    }
    if (deferLast >= 0) {
        w.T[2 * deferLast] = lD1;
        w.T[2 * deferLast + 1] = lB1;
        if (three) {
            w.T[2 * deferLast + 2] = lD2;
            w.T[2 * deferLast + 3] = lB2;
        }
    } else {
        double Dl = three ? D2 : D1;
        double Bl = three ? B2 : B1;
        Dl -= lD1;
        Bl -= lB1;
        if (three) {
            Dl -= lD2;
            Bl -= lB2;
        }
        const int last = three ? m2 : m1;
        w.D[last] = Dl;
        w.B[last] = Bl;
    }
}

// Back-substitution of a unit of two or three members (HSolvePassive.cpp:
// 344-369): VMid of the last member is known; B holds VMid once solved.
// Every value the unit reads is read before it writes (one LDS round trip).
template <bool FD = false>
__device__ __forceinline__ void bwd_small(const Desc& u, Work w) {
    const bool three = (u.a.x & 15) == 3;
    const int m0 = u.a.w & 0xffff, m1 = (u.a.w >> 16) & 0xffff, m2 = u.b.x;
    const double a01 = desc_double(u, 6);
    const double Bm0 = w.B[m0], Dm0 = w.D[m0], Bm1 = w.B[m1];
    double v2 = 0.0, f = 0.0, Dm1 = 1.0;
    if (three) {
        v2 = w.B[m2];
        f = w.F[u.b.y];
        Dm1 = w.D[m1];
    }
    double acc;
    if (three) {
        const double a02 = desc_double(u, 8);
        const double v1 = quot<FD>(Bm1 - v2 * f, Dm1);
        w.B[m1] = v1;
        acc = (Bm0 - v2 * a02) - v1 * a01;
    } else {
        acc = Bm0 - a01 * Bm1;
    }
    w.B[m0] = quot<FD>(acc, Dm0);
}

// start of member p's block in a group's junction coefficients (reference
// HJ_ layout: for each non-last member, (A[c][y], A[y][c]) per later y)
__host__ __device__ constexpr int hb(int S, int p) { return p * (2 * (S - 1) - (p - 1)); }

// Junction groups of four or five members (descriptor form): the members,
// the offset of the group's junction coefficients and the deferred term
// slots come in the descriptor; the group's rows and coefficients are loaded
// into registers at once, eliminated in registers in the operation order of
// the in-memory form below, and stored back -- one memory round trip per
// group instead of one per operation.
template <int S>
__device__ __forceinline__ void group_members(const Desc& u, int (&m)[S]) {
    m[0] = u.a.w & 0xffff;
    m[1] = (u.a.w >> 16) & 0xffff;
    m[2] = u.b.x & 0xffff;
    m[3] = (u.b.x >> 16) & 0xffff;
    if constexpr (S == 5) m[4] = u.b.y;
}

template <int S, bool FD = false>
__device__ __forceinline__ void fwd_group_desc(const Desc& u, const double* hj0, Work w) {
    int m[S];
    group_members<S>(u, m);
    const int deferLast = u.a.y;
    const int hjOff = u.a.z;
    const int dinfo[3] = {u.b.z, u.b.w, u.c.x};   // members 1..3
    double HJ[S * (S - 1)];
#pragma unroll
    for (int k = 0; k < S * (S - 1); ++k) HJ[k] = hj0[hjOff + k];
    double D[S], B[S];
#pragma unroll
    for (int p = 0; p < S; ++p) {
        D[p] = w.D[m[p]];
        B[p] = w.B[m[p]];
    }
#pragma unroll
    for (int p = 0; p + 1 < S; ++p) {
        if (p >= 1 && dinfo[p - 1] >= 0) {
            const int d = dinfo[p - 1] & 0xffffff, cnt = dinfo[p - 1] >> 24;
            for (int t = 0; t < cnt; ++t) {
                D[p] -= w.T[2 * (d + t)];
                B[p] -= w.T[2 * (d + t) + 1];
            }
        }
        const double pivot = D[p];
        const double Bc = B[p];
        const int base = hb(S, p);
#pragma unroll
        for (int q = p + 1; q < S; ++q) {
            const double lower = HJ[base + 2 * (q - p - 1) + 1];
            const double lp = quot<FD>(lower, pivot);
            const double tD = lp * HJ[base + 2 * (q - p - 1)];
            const double tB = lp * Bc;
            if (q == S - 1 && deferLast >= 0) {
                w.T[2 * (deferLast + p)] = tD;
                w.T[2 * (deferLast + p) + 1] = tB;
            } else {
                D[q] -= tD;
                B[q] -= tB;
            }
#pragma unroll
            for (int z = p + 1; z < S; ++z) {
                if (z == q) continue;
                const double above = HJ[base + 2 * (z - p - 1)];
                const int idx = q < z ? hb(S, q) + 2 * (z - q - 1) : hb(S, z) + 2 * (q - z - 1) + 1;
                HJ[idx] -= lp * above;
            }
        }
    }
    // members 1..S-2 always change; the last only without a deferred slot
#pragma unroll
    for (int p = 1; p < S; ++p) {
        if (p == S - 1 && deferLast >= 0) break;
        w.D[m[p]] = D[p];
        w.B[m[p]] = B[p];
    }
#pragma unroll
    for (int k = 0; k < S * (S - 1); ++k) w.HJ[hjOff + k] = HJ[k];
}

template <int S, bool FD = false>
__device__ __forceinline__ void bwd_group_desc(const Desc& u, Work w) {
    int m[S];
    group_members<S>(u, m);
    const int hjOff = u.a.z;
    double HJ[S * (S - 1)];
#pragma unroll
    for (int k = 0; k < S * (S - 1); ++k) HJ[k] = w.HJ[hjOff + k];
    double V[S], D[S];
#pragma unroll
    for (int p = 0; p < S; ++p) {
        V[p] = w.B[m[p]];
        D[p] = w.D[m[p]];
    }
#pragma unroll
    for (int p = S - 2; p >= 0; --p) {
        const int base = hb(S, p);
        double acc = V[p];
#pragma unroll
        for (int q = S - 1; q > p; --q) acc -= V[q] * HJ[base + 2 * (q - p - 1)];
        V[p] = quot<FD>(acc, D[p]);
    }
#pragma unroll
    for (int p = 0; p + 1 < S; ++p) w.B[m[p]] = V[p];
}

// Groups in memory: member lists, deferred slots and the working junction
// coefficients read and written operation by operation.
__device__ __noinline__ void fwd_group_mem(char* arena, const CellDev& c, Desc u, Work w) {
    const int32_t* defer = A_<const int32_t>(arena, c.oDefer);
    const int g = u.a.y;
    const int4 gi = A_<const int4>(arena, c.oGrpInfo)[g];
    const int32_t* mem = A_<const int32_t>(arena, c.oGrpMem) + gi.x;
    const int s = gi.y;
    double* HJ = w.HJ;
    const double* hj0 = w.HJ0;
    for (int k = 0; k < s * (s - 1); ++k) HJ[gi.z + k] = hj0[gi.z + k];
    const int deferLast = gi.w;
    for (int p = 0; p + 1 < s; ++p) {
        const int cn = mem[p];
        if (p >= 1) {
            const int d = defer[2 * cn];
            if (d >= 0) {
                const int cnt = defer[2 * cn + 1];
                for (int t = 0; t < cnt; ++t) {
                    w.D[cn] -= w.T[2 * (d + t)];
                    w.B[cn] -= w.T[2 * (d + t) + 1];
                }
            }
        }
        const double pivot = w.D[cn];
        const double Bc = w.B[cn];
        const int base = gi.z + hb(s, p);
        for (int q = p + 1; q < s; ++q) {
            const int y = mem[q];
            const double lower = HJ[base + 2 * (q - p - 1) + 1];
            const double tD = lower / pivot * HJ[base + 2 * (q - p - 1)];
            const double tB = lower / pivot * Bc;
            if (q == s - 1 && deferLast >= 0) {
                w.T[2 * (deferLast + p)] = tD;
                w.T[2 * (deferLast + p) + 1] = tB;
            } else {
                w.D[y] -= tD;
                w.B[y] -= tB;
            }
            for (int z = p + 1; z < s; ++z) {
                if (z == q) continue;
                const double above = HJ[base + 2 * (z - p - 1)];
                const int idx = q < z ? gi.z + hb(s, q) + 2 * (z - q - 1) : gi.z + hb(s, z) + 2 * (q - z - 1) + 1;
                HJ[idx] -= lower / pivot * above;
            }
        }
    }
}

__device__ __noinline__ void bwd_group_mem(char* arena, const CellDev& c, Desc u, Work w) {
    const int g = u.a.y;
    const int4 gi = A_<const int4>(arena, c.oGrpInfo)[g];
    const int32_t* mem = A_<const int32_t>(arena, c.oGrpMem) + gi.x;
    const int s = gi.y;
    for (int p = s - 2; p >= 0; --p) {
        const int cn = mem[p];
        const int base = gi.z + hb(s, p);
        double acc = w.B[cn];
        for (int q = s - 1; q > p; --q) acc -= w.B[mem[q]] * w.HJ[base + 2 * (q - p - 1)];
        w.B[cn] = acc / w.D[cn];
    }
}

template <bool SLOWG, bool FD = false>
__device__ __forceinline__ void forward_unit(char* arena, const CellDev& c, const Desc& u, Work w) {
    if constexpr (SLOWG) {
        if (u.a.x & 32) {
            // inlined, so that their row and term accesses stay LDS
            // instructions (through a call they become flat ones, each
            // waiting for every global load in flight)
            if ((u.a.x & 15) == 4) fwd_group_desc<4, FD>(u, w.HJ0, w);
            else fwd_group_desc<5, FD>(u, w.HJ0, w);
            return;
        }
        if (u.a.x & 16) {
            fwd_group_mem(arena, c, u, w);
            return;
        }
    }
    fwd_small<FD>(u, w);
}

template <bool SLOWG, bool FD = false>
__device__ __forceinline__ void backward_unit(char* arena, const CellDev& c, const Desc& u, Work w) {
    if constexpr (SLOWG) {
        if (u.a.x & 32) {
            if ((u.a.x & 15) == 4) bwd_group_desc<4, FD>(u, w);
            else bwd_group_desc<5, FD>(u, w);
            return;
        }
        if (u.a.x & 16) {
            bwd_group_mem(arena, c, u, w);
            return;
        }
    }
    bwd_small<FD>(u, w);
}

// One level's units [s, e) in one direction: unit s + tid from `d` (loaded
// ahead by the caller), later ones of a level wider than the workgroup as
// they come.
template <bool FWD, bool SLOWG, bool FD>
__device__ __forceinline__ void sweep_level(char* arena, const CellDev& c, const int4* __restrict__ desc,
                                            const Desc& d, int s, int e, Work w) {
    const int tid = threadIdx.x, nt = blockDim.x;
    if (s + tid < e) {
        if constexpr (FWD) forward_unit<SLOWG, FD>(arena, c, d, w);
        else backward_unit<SLOWG, FD>(arena, c, d, w);
    }
    for (int u = s + tid + nt; u < e; u += nt) {
        if constexpr (FWD) forward_unit<SLOWG, FD>(arena, c, load_desc(desc, u), w);
        else backward_unit<SLOWG, FD>(arena, c, load_desc(desc, u), w);
    }
}

// CaConcStruct::process (hsolve/HSolveStruct.cpp:138-155)
__device__ __forceinline__ double caconc_process(double& cst, double basal, double f1, double f2,
                                                double ceil, double floor, double act) {
    cst = f1 * cst + f2 * act;
    double ca = basal + cst;
    if (ceil > 0 && ca > ceil) {
        ca = ceil;
        cst = ca - basal;
    }
    if (ca < floor) {
        ca = floor;
        cst = ca - basal;
    }
    return ca;
}

// advanceCalcium for the pools of compartment i (HSolveActive.cpp:153-216):
// activation = iCa added between steps, then the channels feeding the pool
// in channel order.  vmid: VMid of the compartment, vnew: its new Vm.  Every
// load of a pool is issued before its stores, so they travel together.
template <typename GT>
__device__ __forceinline__ void calcium_row(char* arena, const CellDev& c, const View<GT>& v,
                                            const ChanType* __restrict__ types, int i, double vmid, double vnew,
                                            bool firstOfCall) {
    const int n = v.n;
    const int nPools = v.poolCount[i];
    double* ica = A_<double>(arena, c.oPoolICa);
    double* pc = A_<double>(arena, c.oPoolC);
    for (int q = 0; q < nPools; ++q) {
        const size_t o = (size_t)q * n + i;
        double cst = pc[o];
        const double basal = A_<const double>(arena, c.oPoolBasal)[o];
        const double f1 = A_<const double>(arena, c.oPoolF1)[o];
        const double f2 = A_<const double>(arena, c.oPoolF2)[o];
        const double ceil = A_<const double>(arena, c.oPoolCeil)[o];
        const double floor = A_<const double>(arena, c.oPoolFloor)[o];
        double act = 0.0;
        if (firstOfCall) act = 0.0 + ica[o];
        int js = 0;
        for (int k = 0; k < v.nChanSlots; ++k) {
            const int t = v.uniform ? k : v.chType[(size_t)k * n + i];
            if (t < 0) continue;
            const ChanType& ct = types[t];
            if (v.uniform && ct.nGates < 0) continue;
            if (ct.caTgtLocal == q) {
                GT fraction = v.hasModulation ? (GT)v.mod[(size_t)k * n + i] : (GT)1.0;
                int jj = js;
                if (ct.Xpower > 0.0) fraction *= take_power<GT>((GT)v.state[(size_t)(jj++) * n + i], ct.Xpower);
                if (ct.Ypower > 0.0) fraction *= take_power<GT>((GT)v.state[(size_t)(jj++) * n + i], ct.Ypower);
                if (ct.Zpower > 0.0) fraction *= take_power<GT>((GT)v.state[(size_t)(jj++) * n + i], ct.Zpower);
                const double Gk = (double)(v.Gbar[(size_t)k * n + i] * fraction);
                const double ek = (double)v.Ek[(size_t)k * n + i];
                if (v.caAdvance == 1) act += Gk * (ek - vmid);
                else act += Gk * (ek - (2 * vmid - vnew));
            }
            js += ct.nGates;
        }
        v.poolCa[o] = caconc_process(cst, basal, f1, f2, ceil, floor, act);
        pc[o] = cst;
        if (firstOfCall) ica[o] = 0.0;
    }
}

// advanceCalcium (calcium_row's operations, in its order) for a cell whose
// four channel slots are Na, Kdr, KA and CaL (kSigCa4: gate states at 0-1,
// 2, 3-4 and 5; powers 3/1, 4, 3/1 and 2; no modulation) with at most one
// pool per compartment.  `feed`: bit k set when slot k's channel feeds the
// pool (from the staged types; the same for every compartment).  The loads
// of a compartment (its old Vm, the pool's inputs, and the states, Gbar and
// Ek of the feeding slots) are issued together by ca4_load; ca4_finish does
// the arithmetic and the stores.
template <typename GT>
struct Ca4In {
    double V, pc, basal, f1, f2, ceil, floor, ica;
    double st[6];
    GT gb[4], ek[4];
    int np;
};

template <typename GT>
__device__ __forceinline__ void ca4_load(char* arena, const CellDev& c, const View<GT>& v, int feed, int i,
                                         bool first, Ca4In<GT>& R) {
    constexpr int base[4] = {0, 2, 3, 5}, ng[4] = {2, 1, 2, 1};
    const int n = v.n;
    R.V = v.V[i];
    R.np = v.poolCount[i];
    R.pc = A_<const double>(arena, c.oPoolC)[i];
    R.basal = A_<const double>(arena, c.oPoolBasal)[i];
    R.f1 = A_<const double>(arena, c.oPoolF1)[i];
    R.f2 = A_<const double>(arena, c.oPoolF2)[i];
    R.ceil = A_<const double>(arena, c.oPoolCeil)[i];
    R.floor = A_<const double>(arena, c.oPoolFloor)[i];
    R.ica = first ? A_<const double>(arena, c.oPoolICa)[i] : 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!((feed >> k) & 1)) continue;
#pragma unroll
        for (int g = 0; g < ng[k]; ++g) R.st[base[k] + g] = v.state[(size_t)(base[k] + g) * n + i];
        R.gb[k] = v.Gbar[(size_t)k * n + i];
        R.ek[k] = v.Ek[(size_t)k * n + i];
    }
}

template <typename GT>
__device__ __forceinline__ void ca4_finish(char* arena, const CellDev& c, const View<GT>& v, int feed, int i,
                                           double vmid, bool first, Ca4In<GT>& R) {
    constexpr int base[4] = {0, 2, 3, 5}, ng[4] = {2, 1, 2, 1};
    constexpr double pw[4][2] = {{3.0, 1.0}, {4.0, 0.0}, {3.0, 1.0}, {2.0, 0.0}};
    const double vnew = 2 * vmid - R.V;
    v.V[i] = vnew;
    if (R.np < 1) return;
    double act = 0.0;
    if (first) act = 0.0 + R.ica;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!((feed >> k) & 1)) continue;
        GT fraction = (GT)1.0;
#pragma unroll
        for (int g = 0; g < ng[k]; ++g) fraction *= take_power<GT>((GT)R.st[base[k] + g], pw[k][g]);
        const double Gk = (double)(R.gb[k] * fraction);
        const double ek = (double)R.ek[k];
        if (v.caAdvance == 1) act += Gk * (ek - vmid);
        else act += Gk * (ek - (2 * vmid - vnew));
    }
    double cst = R.pc;
    v.poolCa[i] = caconc_process(cst, R.basal, R.f1, R.f2, R.ceil, R.floor, act);
    A_<double>(arena, c.oPoolC)[i] = cst;
    if (first) A_<double>(arena, c.oPoolICa)[i] = 0.0;
}

// Copies the slot types of a uniform cell to LDS (empty slots get
// nGates = -1).  Callers then read a uniform cell's types from sTypes and
// other cells' from the global table, each through its own pointer, so the
// compiler knows the address space of either (a pointer that may be either
// makes every read a flat access, which waits for every load in flight).
__device__ __forceinline__ void stage_types(const CellDev& c, const ChanType* types, ChanType* sTypes) {
    if (!c.uniform) return;
    for (int k = threadIdx.x; k < c.nChanSlots; k += blockDim.x) {
        const int t = c.slotType[k];
        if (t >= 0) sTypes[k] = types[t];
        else sTypes[k].nGates = -1;
    }
}

// PulseGen's period (device/PulseGen.cpp:362-369).
__device__ __forceinline__ double pulse_period(const PulseDev& p) {
    double period = p.width[0] + p.delay[0];
    for (int ii = 1; ii < p.n && (p.width[ii] > 0.0 || p.delay[ii] > 0.0); ++ii) {
        const double incr = p.delay[ii] + p.width[ii] - p.width[ii - 1];
        if (incr > 0) period += incr;
    }
    return period;
}

// The level for a phase inside the period (device/PulseGen.cpp:408-421);
// `out` where no pulse or gap holds the phase (the reference keeps output_).
__device__ __forceinline__ double pulse_level(const PulseDev& p, double phase, double out) {
    for (int ii = 0; ii < p.n; ++ii) {
        if (phase < p.delay[ii]) return p.base;
        if (phase < p.delay[ii] + p.width[ii]) return p.level[ii];
        phase -= p.delay[ii];
    }
    return out;
}

// The output at time t of a free-running PulseGen that drives another's
// input: a function of t (the setup admits only sources whose pulses and
// gaps cover the period, so the level is always set).
__device__ __forceinline__ double pulse_free_output(const PulseDev& s, double t) {
    return pulse_level(s, fmod(t, pulse_period(s)), s.base);
}

// PulseGen::process (device/PulseGen.cpp:359-424) at time t with input
// `input` (trigger modes): updates the output and trigger state; false
// when the reference sends nothing (phase past the period).
__device__ __noinline__ bool pulse_tick(PulseDev& p, double t, double input) {
    const double period = pulse_period(p);
    double phase;
    if (p.trigMode == 0) {
        phase = fmod(t, period);
    } else if (p.trigMode == 1) {          // EXT_TRIG
        if (input == 0) {
            phase = p.trigTime < 0 ? period : t - p.trigTime;
        } else {
            if (p.prevInput == 0) p.trigTime = t;
            phase = t - p.trigTime;
        }
        p.prevInput = input;
    } else {                               // EXT_GATE
        if (input == 0) {
            phase = period;
        } else {
            if (p.prevInput == 0) p.trigTime = t;
            phase = fmod(t - p.trigTime, period);
        }
        p.prevInput = input;
    }
    if (phase >= period) {
        p.out = p.base;
        return false;
    }
    p.out = pulse_level(p, phase, p.out);
    return true;
}

// VClamp::process (biophysics/VClamp.cpp:287-311) with the sensed Vm vIn
// (what the solver sent at the end of the previous tick): the low-pass
// filtered command, the error and the PID current increment; returns the
// current it sends.
__device__ __noinline__ double vclamp_tick(VClampDev& c, double vIn) {
    const double dCmd = c.cmdIn - c.oldCmdIn;
    c.command = c.cmdIn + dCmd * (1 - c.tauByDt) + (c.command - c.cmdIn + dCmd * c.tauByDt) * c.expt;
    c.oldCmdIn = c.cmdIn;
    const double e = c.command - vIn;
    if (c.mode == 0) {
        c.current += c.Kp * ((1 + c.dtByTi + c.tdByDt) * e - (1 + 2 * c.tdByDt) * c.e1 + c.tdByDt * c.e2);
        c.e2 = c.e1;
        c.e1 = e;
    } else if (c.mode == 1) {
        c.current += c.Kp * ((1 + c.dtByTi) * e - c.e1 + c.tdByDt * (vIn - 2 * c.v1 + c.e2));
        c.e2 = c.v1;
        c.v1 = vIn;
        c.e1 = e;
    } else {
        c.current += c.Kp * (vIn - c.v1 + c.dtByTi * e + c.tdByDt * (vIn - 2 * c.v1 + c.e2));
        c.e2 = c.v1;
        c.v1 = vIn;
    }
    return c.current;
}

// The varying injection of entry e for the tick ending at t (tick stepNo,
// 1 the first after reinit): the call's host additions (first tick), then
// what the entry's PulseGens and VClamps send (clock tick 0), in their order
// (HSolve::addInject sums into injectVarying).  A PulseGen driven by another
// sees that one's output as the reference's immediate message delivers it:
// this tick's when the source is processed first, else the previous tick's
// (its reinit baseLevel at the first tick).
// The host's inputs of injection entry e (addInject since the last call),
// taken on the call's first tick.
__device__ __forceinline__ double inject_host(char* arena, const CellDev& c, int e, bool first) {
    if (!first) return 0.0;
    double* iv = A_<double>(arena, c.oInjVar);
    const double var = iv[e];
    iv[e] = 0.0;
    return var;
}

__device__ __forceinline__ double inject_varying(char* arena, const CellDev& c, int e, bool first, double t,
                                                 double dt, int64_t stepNo) {
    double* iv = A_<double>(arena, c.oInjVar);
    double var = first ? iv[e] : 0.0;
    if (first) iv[e] = 0.0;
    if (e >= c.nSetupEntries) return var;   // an entry added later: host inputs only
    if (c.nPulse > 0) {
        PulseDev* pl = A_<PulseDev>(arena, c.oPulse);
        const int32_t* st = A_<const int32_t>(arena, c.oPulseStart);
        const int32_t* ix = A_<const int32_t>(arena, c.oPulseIdx);
        for (int q = st[e]; q < st[e + 1]; ++q) {
            PulseDev& p = pl[ix[q]];
            double in = p.input;
            if (p.src >= 0) {
                const PulseDev& sp = pl[p.src];
                in = p.srcBefore ? pulse_free_output(sp, t)
                     : stepNo <= 1 ? sp.base : pulse_free_output(sp, dt * (double)(unsigned long long)(stepNo - 1));
            }
            if (pulse_tick(p, t, in)) var += p.out;
            // input_ as read after the tick: the source's last message
            if (p.src >= 0) p.input = p.srcBefore ? in : pulse_free_output(pl[p.src], t);
        }
    }
    if (c.nVClamp > 0) {
        VClampDev* vc = A_<VClampDev>(arena, c.oVClamp);
        const int32_t* st = A_<const int32_t>(arena, c.oVcStart);
        const int32_t* ix = A_<const int32_t>(arena, c.oVcIdx);
        const double* V = A_<const double>(arena, c.oV);
        for (int q = st[e]; q < st[e + 1]; ++q) var += vclamp_tick(vc[ix[q]], V[vc[ix[q]].compt]);
    }
    return var;
}

// The probe schedule of one launch, worked out once at its start: probes
// sample every probeStride-th step (a Table on a slower clock tick,
// Table::process after the solver's tick), sample 0 being the first such
// step after the trace's base step.  `next` is the next sampled step and
// `sample` its index in the trace.
struct ProbeClock {
    int64_t next, stride, sample;
};

__device__ __forceinline__ ProbeClock probe_clock(char* arena, const CellDev& c, int64_t step0) {
    ProbeClock p{-1, 1, 0};
    if (c.nProbe <= 0) return p;
    const int64_t base = *A_<const int64_t>(arena, c.oProbeBase);
    p.stride = c.probeStride;
    p.next = (step0 / p.stride + 1) * p.stride;   // the first multiple of the stride after step0
    p.sample = p.next / p.stride - base / p.stride - 1;
    return p;
}

// SpikeGen::process at t = dt * step for generator sg, given Vm
// (biophysics/SpikeGen.cpp:202-216): its last spike time and fired flag are
// read and written by the caller.
__device__ __forceinline__ void spike_rule(const LaunchArgs& a, const CellDev& c, int sg, double Vm, double thr,
                                          double refr, int edge, double& last, int& fired, int64_t stepNo) {
    if (Vm > thr) {
        const double t = a.dt * (double)(unsigned long long)stepNo;
        if ((t + a.dt / 2.0) >= (last + refr)) {
            if (!(edge && fired)) {
                const unsigned long long slot = atomicAdd(a.spikeCount, 1ULL);
                if (slot < (unsigned long long)a.spikeCap) {
                    SpikeRec r;
                    r.cell = c.cellId;
                    r.gen = sg;
                    r.step = stepNo;
                    a.spikes[slot] = r;
                }
                last = t;
                fired = 1;
            }
        }
    } else {
        fired = 0;
    }
}

// Per-tick outputs once V of the tick is complete: spike generators
// (SpikeGen::process at t = dt * step, biophysics/SpikeGen.cpp:202-216, fed
// by HSolveActive::sendSpikes, hsolve/HSolveActive.cpp:342-347) into the
// engine's spike ring, and Vm probes (Table sampling after the solver's tick).
__device__ __forceinline__ void tick_outputs(const LaunchArgs& a, const CellDev& c, const double* __restrict__ V,
                                             int64_t stepNo, ProbeClock& pc) {
    char* arena = a.arena;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int sg = tid; sg < c.nSpike; sg += nt) {
        const int comp = A_<const int32_t>(arena, c.oSpkCompt)[sg];
        const double Vm = V[comp];
        const double thr = A_<const double>(arena, c.oSpkThr)[sg];
        int32_t* firedP = A_<int32_t>(arena, c.oSpkFired);
        double* lastP = A_<double>(arena, c.oSpkLast);
        int fired = firedP[sg];
        double last = lastP[sg];
        const int fired0 = fired;
        const double last0 = last;
        spike_rule(a, c, sg, Vm, thr, A_<const double>(arena, c.oSpkRefr)[sg], A_<const int32_t>(arena, c.oSpkEdge)[sg],
                   last, fired, stepNo);
        if (fired != fired0) firedP[sg] = fired;
        if (last != last0) lastP[sg] = last;
    }
    // probes (ProbeClock)
    if (c.nProbe > 0 && stepNo == pc.next) {
        const int64_t sample = pc.sample;
        if (sample >= 0 && sample < c.probeCap) {
            const int64_t* poff = A_<const int64_t>(arena, c.oProbeOff);
            double* tr = A_<double>(arena, c.oProbeTrace);
            for (int j = tid; j < c.nProbe; j += nt) tr[sample * c.nProbe + j] = *A_<const double>(arena, poff[j]);
        }
        pc.next += pc.stride;
        ++pc.sample;
    }
}

// ---------------------------------------------------------------------------
// the step kernel
// ---------------------------------------------------------------------------
// K ticks of one cell (one workgroup).  SPLIT: phase A ran in
// hsolve_channel_kernel, which left the matrix rows in the cell's HBM work
// arrays; the workgroup only stages them into LDS.  WL: the elimination work
// arrays and the level table live in LDS (else in the cell's HBM work area);
// a compile-time choice so that every access is a plain LDS or global
// instruction rather than a generic (flat) one.
// The channel pass of a uniform cell with integer powers (every compartment
// has the channel types slotType[], at most MS of them, packed in
// CellDev::fast and read with scalar loads, so every branch on them is
// uniform), with every global load of the compartment issued before its
// arithmetic: gate states, Gbar, Ek, modulation, Ca, V and the passive terms
// first, then the table rows of all gates (no load waits on a store: the new
// states and the row are stored last).  The arithmetic is channel_row's, in
// the same order.
template <typename GT>
__device__ __forceinline__ GT power_code(GT x, int code) {
    if (code == 1) return x;
    if (code == 2) return x * x;
    if (code == 3) return x * x * x;
    GT y = x * x;
    return y * y;
}

// V is the compartment's Vm, loaded by the caller (the fused kernel loads it
// one compartment ahead, so that the table rows need no load of their own).
// NOPOOL: the caller knows the cell has no Ca pools.  SIG > 0: the caller
// knows the cell's channel slots (no Z gate, no instant gate, no
// modulation), so the gate codes are constants and every branch on them
// folds away; only the table columns and state slots are read from the cell.
// kSigHH: the squid HH pair (m^3 h, then n^4); kSigCa4: m^3 h, n^4, a^3 b,
// s^2 (Na, Kdr, KA, CaL).
constexpr int kSigHH = 1;
constexpr int kSigCa4 = 2;
constexpr int kSigSlots[3] = {0, 2, 4};
constexpr int kSigWord[3][4] = {{0, 0, 0, 0}, {3 | 1 << 3, 4, 0, 0}, {3 | 1 << 3, 4, 3 | 1 << 3, 2}};

// The loaded inputs of one compartment's fast channel pass.
template <typename GT, int MS>
struct RowIn {
    double Dbase, CmByDt, EmByRm;
    double st[MS][3];
    GT gb[MS], ek[MS];
    double md[MS];
    GT raw[MS][3][4];
    GT zf[MS];
    GT frac;
};

// Every global load of compartment i's channel pass, issued together: gate
// states, Gbar, Ek, modulation, Ca inputs, the passive terms, then the table
// entries of every gate ((A, B) at the row and the next row; Z: the Ca row of
// its pool, the channel's external Ca, or V, as channel_row chooses).
// Absent entries read valid addresses and are not used.
template <typename GT, int MS, bool NOPOOL, int SIG>
__device__ __forceinline__ void row_load(const View<GT>& v, const CellDev& c, int i, const double V,
                                         RowIn<GT, MS>& R) {
    static_assert(SIG == 0 || MS == kSigSlots[SIG], "a signature fixes the slot count");
    const int n = v.n;
    const int S = SIG ? kSigSlots[SIG] : v.nChanSlots;
    auto word2 = [&](int k) { return SIG ? kSigWord[SIG][k] : c.fast[k][2]; };
    auto live = [&](int k) { return SIG ? k < kSigSlots[SIG] : (k < S && c.fast[k][1] < 0); };
    const bool hasPools = !NOPOOL && v.hasPools;
    R.Dbase = v.Dbase[i];
    R.CmByDt = v.CmByDt[i];
    R.EmByRm = v.EmByRm[i];
    const int nPools = hasPools ? v.poolCount[i] : 0;
    const double* __restrict__ caSrc = hasPools ? v.poolCa : v.V;
    const double* __restrict__ stSrc = c.nGateSlots > 0 ? v.state : v.V;
    double cav[MS], ext[MS];
#pragma unroll
    for (int k = 0; k < MS; ++k) {
        const int w1 = c.fast[k][1], w2 = word2(k);
        const size_t kk = (size_t)(k < S ? k : S - 1) * n + i;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
            R.st[k][ax] = SIG == 0 || ((w2 >> (3 * ax)) & 7) != 0
                              ? stSrc[(size_t)((w1 >> (16 + 5 * ax)) & 31) * n + i] : 0.0;
        R.gb[k] = v.Gbar[kk];
        R.ek[k] = v.Ek[kk];
        R.md[k] = SIG == 0 && v.hasModulation ? v.mod[kk] : 1.0;
        // Ca inputs only where a Z gate may read them (uniform condition)
        const bool hasZ = ((w2 >> 6) & 7) != 0;
        ext[k] = hasZ ? v.extCa[kk] : 0.0;
        const int caDep = ((w2 >> 12) & 15) - 1;
        cav[k] = hasZ && caDep >= 0 ? caSrc[(size_t)(hasPools && caDep > 0 ? caDep : 0) * n + i] : 0.0;
    }
    int vBase;
    double vFrac;
    table_row(v.vt, V, vBase, vFrac);
    R.frac = (GT)vFrac;
#pragma unroll
    for (int k = 0; k < MS; ++k) {
        const int w0 = c.fast[k][0], w1 = c.fast[k][1], w2 = word2(k);
        if (!live(k)) continue;                   // empty slot (live bit 31 clear)
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            if (((w2 >> (3 * ax)) & 7) == 0) continue;
            const int col = ax == 0 ? (w0 & 0xffff) : ax == 1 ? ((w0 >> 16) & 0xffff) : (w1 & 0xffff);
            const GT* tab = v.vt.tab;
            int base = vBase, nCol = v.vt.nCol;
            if (ax == 2) {
                R.zf[k] = R.frac;
                const int caDep = ((w2 >> 12) & 15) - 1;
                const bool fromPool = caDep >= 0 && caDep < nPools;
                const bool fromExt = !fromPool && ext[k] > 0 && col + 1 < v.ct.nCol;
                if (fromPool || fromExt) {
                    double cf;
                    table_row(v.ct, fromPool ? cav[k] : ext[k], base, cf);
                    R.zf[k] = (GT)cf;
                    tab = v.ct.tab;
                    nCol = v.ct.nCol;
                }
            }
            const GT* ap = tab + base + col;
            R.raw[k][ax][0] = ap[0];
            R.raw[k][ax][1] = ap[nCol];
            R.raw[k][ax][2] = ap[1];
            R.raw[k][ax][3] = ap[nCol + 1];
        }
    }
}

// The arithmetic of compartment i's channel pass (channel_row's, in the same
// order) and its stores: the new gate states, and the row (D, B) returned.
template <typename GT, int MS, bool NOPOOL, int SIG>
__device__ __forceinline__ void row_finish(const View<GT>& v, const CellDev& c, int i, const double V, double dtg,
                                           RowIn<GT, MS>& R, double& Dout, double& Bout) {
    const int n = v.n;
    const int S = SIG ? kSigSlots[SIG] : v.nChanSlots;
    auto word2 = [&](int k) { return SIG ? kSigWord[SIG][k] : c.fast[k][2]; };
    auto live = [&](int k) { return SIG ? k < kSigSlots[SIG] : (k < S && c.fast[k][1] < 0); };
    double GkSum = 0.0, GkEkSum = 0.0;
#pragma unroll
    for (int k = 0; k < MS; ++k) {
        const int w2 = word2(k);
        if (!live(k)) continue;
        GT fraction = SIG == 0 && v.hasModulation ? (GT)R.md[k] : (GT)1.0;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int code = (w2 >> (3 * ax)) & 7;
            if (code == 0) continue;
            // LookupTable::lookup (RateLookup.cpp:78-103)
            const GT fr = ax == 2 ? R.zf[k] : R.frac;
            const GT C1 = R.raw[k][ax][0] + (R.raw[k][ax][1] - R.raw[k][ax][0]) * fr;
            const GT C2 = R.raw[k][ax][2] + (R.raw[k][ax][3] - R.raw[k][ax][2]) * fr;
            const double x = gate_step<std::is_same<GT, float>::value>(R.st[k][ax], C1, C2, dtg, (w2 >> (9 + ax)) & 1);
            R.st[k][ax] = x;
            fraction *= power_code<GT>((GT)x, code);
        }
        const GT Gk = R.gb[k] * fraction;
        GkSum += (double)Gk;
        GkEkSum += (double)Gk * (double)R.ek[k];
    }
#pragma unroll
    for (int k = 0; k < MS; ++k) {
        const int w1 = c.fast[k][1], w2 = word2(k);
        if (!live(k)) continue;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
            if (((w2 >> (3 * ax)) & 7) != 0) v.state[(size_t)((w1 >> (16 + 5 * ax)) & 31) * n + i] = R.st[k][ax];
    }
    Dout = R.Dbase + GkSum;
    Bout = V * R.CmByDt + R.EmByRm + GkEkSum;
}

template <typename GT, int MS, bool NOPOOL = false, int SIG = 0>
__device__ __forceinline__ void channel_row_fast(const View<GT>& v, const CellDev& c, int i, const double V,
                                                 double dtg, double& Dout, double& Bout) {
    RowIn<GT, MS> R;
    row_load<GT, MS, NOPOOL, SIG>(v, c, i, V, R);
    row_finish<GT, MS, NOPOOL, SIG>(v, c, i, V, dtg, R, Dout, Bout);
}

// FASTMS > 0: the cells are uniform with at most FASTMS channel slots and
// phase A uses channel_row_fast.  LEAN: the cells have no PulseGen, VClamp,
// HHChannel2D, Leakage or SynChan objects and no Ca pools, so those paths
// (and the registers they hold) are left out.
template <typename GT, bool SPLIT, bool WL, bool SLOWG, int FASTMS, bool LEAN = false, int SIG = 0>
__device__ __forceinline__ void cell_steps(const LaunchArgs& a, const CellDev& c, double* lds, ChanType* sTypes) {
    char* arena = a.arena;
    const View<GT> v = make_view<GT>(arena, c, a.tsets);
    const int n = v.n;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    // phase D's old Vm of four compartments per thread loaded ahead (the
    // mixed HH lean kernel, whose registers leave room for it: eight spill)
    constexpr bool VPRE = LEAN && SIG == kSigHH && std::is_same<GT, float>::value;
    constexpr int kVPre = VPRE ? 4 : 1;
    const int nTerm = c.nTerm;
    const int nLevels = c.nLevels;
    const int nUnits = c.nUnits;
    const int nInj = c.nInj;
    const int nExt = c.nExt;
    const int nSpike = c.nSpike;
    const int nProbe = c.nProbe;
    // work arrays: D[n], B[n], T[2*nTerm], F[nFill]; HJ (slow groups, HBM)
    const int nWork = 2 * n + 2 * nTerm + c.nFill;
    double* gwork = A_<double>(arena, c.oWork);
    Work w;
    if constexpr (WL) {
        w.D = lds;
        w.B = lds + n;
        w.T = lds + 2 * n;
        w.F = lds + 2 * n + 2 * nTerm;
    } else {
        w.D = gwork;
        w.B = gwork + n;
        w.T = gwork + 2 * n;
        w.F = gwork + 2 * n + 2 * nTerm;
    }
    w.HJ = gwork + nWork;
    w.HJ0 = A_<const double>(arena, c.oHJ0);
    // the groups' junction coefficients in LDS (hjInLds): staged from their
    // initial values at the start of every tick and worked on in place
    const bool hjLds = SLOWG && WL && c.hjInLds;
    const int nHJL = hjLds ? c.nHJ : 0;
    const double* gHJ0 = w.HJ0;
    if (hjLds) {
        w.HJ = lds + nWork;
        w.HJ0 = w.HJ;
    }
    // level table: in LDS after the work arrays, or read from HBM
    const int32_t* gLevel = A_<const int32_t>(arena, c.oLevelStart);
    const int32_t* level;
    // the chain: levels [Lc, nLevels) of one unit each (the trunk next to
    // the root), eliminated and back-substituted by thread 0 alone, one unit
    // after another with no barrier between them; its descriptors in LDS,
    // 16-byte aligned after the level table (work arrays in LDS only)
    const int Lc = WL ? c.chainLevel : nLevels;
    const int nChain = nLevels - Lc;
    const int4* sChain = nullptr;
    if constexpr (WL) {
        int32_t* sLevel = reinterpret_cast<int32_t*>(lds + nWork + nHJL);
        for (int l = tid; l <= nLevels; l += nt) sLevel[l] = gLevel[l];
        level = sLevel;
        const size_t off = ((size_t)8 * (nWork + nHJL) + 4 * ((size_t)nLevels + 1) + 15) & ~(size_t)15;
        int4* sc = reinterpret_cast<int4*>(reinterpret_cast<char*>(lds) + off);
        const int4* gd = A_<const int4>(arena, c.oDesc) + (nChain > 0 ? 3 * (size_t)gLevel[Lc] : 0);
        for (int k = tid; k < 3 * nChain; k += nt) sc[k] = gd[k];
        sChain = sc;
    } else {
        level = gLevel;
    }
    stage_types(c, a.types, sTypes);
    const double dtg = a.dt;
    const int4* desc = A_<const int4>(arena, c.oDesc);
    int64_t* stepCounter = A_<int64_t>(arena, c.oSteps);
    const int64_t step0 = *stepCounter;
    const bool uniform = v.uniform;
    // the mixed form divides with quot<true> in the elimination
    constexpr bool FD = std::is_same<GT, float>::value;
    ProbeClock pc = probe_clock(arena, c, step0);
    __syncthreads();
    // kSigCa4 cells (uniform: their types staged in LDS): the slots whose
    // channel feeds the compartment's pool
    int ca4Feed = 0;
    if constexpr (SIG == kSigCa4)
        for (int q = 0; q < 4; ++q) ca4Feed |= (sTypes[q].caTgtLocal == 0 ? 1 : 0) << q;

    for (int k = 0; k < a.nsteps; ++k) {
        const bool first = a.firstOfCall && k == 0;
        // the junction coefficients' initial values, sixteen loads in flight
        // per thread before their stores (the forward sweep reads them after
        // the barrier that follows phase A)
        if (hjLds)
            for (int k0 = tid; k0 < nHJL; k0 += 16 * nt) {
                double hv[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) hv[j] = gHJ0[min(k0 + j * nt, nHJL - 1)];
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (k0 + j * nt < nHJL) w.HJ[k0 + j * nt] = hv[j];
            }
        // first unit of level 0 for this thread, loaded while phase A runs
        Desc cur{};
        cur = load_desc(desc, min(tid, max(nUnits - 1, 0)));
        // ---- A: channels, currents, matrix rows
        if constexpr (SPLIT) {
            if constexpr (WL)
                for (int i = tid; i < n; i += nt) {
                    w.D[i] = gwork[i];
                    w.B[i] = gwork[n + i];
                }
        } else if constexpr (FASTMS > 0) {
            if constexpr (SIG == kSigHH) {
                // two compartments per iteration: both compartments' loads
                // travel together, then both are computed and stored (Vm
                // loaded an iteration ahead, so the table rows need no wait)
                double va = tid < n ? v.V[tid] : 0.0, vb = tid + nt < n ? v.V[tid + nt] : 0.0;
#pragma unroll 1
                for (int i = tid; i < n; i += 2 * nt) {
                    const int j = min(i + nt, n - 1);
                    const double Va = va, Vb = vb;
                    va = v.V[min(i + 2 * nt, n - 1)];
                    vb = v.V[min(i + 3 * nt, n - 1)];
                    RowIn<GT, FASTMS> Ra, Rb;
                    row_load<GT, FASTMS, LEAN, SIG>(v, c, i, Va, Ra);
                    row_load<GT, FASTMS, LEAN, SIG>(v, c, j, Vb, Rb);
                    double Da, Ba;
                    row_finish<GT, FASTMS, LEAN, SIG>(v, c, i, Va, dtg, Ra, Da, Ba);
                    w.D[i] = Da;
                    w.B[i] = Ba;
                    if (i + nt < n) {
                        double Db, Bb;
                        row_finish<GT, FASTMS, LEAN, SIG>(v, c, j, Vb, dtg, Rb, Db, Bb);
                        w.D[j] = Db;
                        w.B[j] = Bb;
                    }
                }
            } else {
                // Vm one compartment ahead: the table rows of a compartment
                // are known when its loads are issued, so they travel with them
                double vNext = tid < n ? v.V[tid] : 0.0;
#pragma unroll 1
                for (int i = tid; i < n; i += nt) {
                    const double Vi = vNext;
                    if (i + nt < n) vNext = v.V[i + nt];
                    channel_row_fast<GT, FASTMS, LEAN, SIG>(v, c, i, Vi, dtg, w.D[i], w.B[i]);
                }
            }
        } else {
            if (uniform)
                for (int i = tid; i < n; i += nt) channel_row<GT, true>(v, sTypes, i, dtg, w, false, i);
            else
                for (int i = tid; i < n; i += nt) channel_row<GT, false>(v, a.types, i, dtg, w, false, i);
        }
        // (split mode: the channel kernel has added them to the rows)
        if (!SPLIT && (nInj > 0 || (first && nExt > 0) || (!LEAN && c.nXObj > 0))) {
            // VClamps read V written by other threads in the last tick: a
            // full barrier then (its release covers those global stores)
            if (!LEAN && c.nVClamp > 0) __syncthreads();
            else work_barrier(WL);
            const int32_t* xflag = A_<const int32_t>(arena, c.oXFlag);
            // injections: HS[4i+3] += injectVarying + injectBasal
            // (HSolveActive.cpp:119-128); one entry per compartment, so
            // the entries are independent
            {
                const int32_t* ii = A_<const int32_t>(arena, c.oInjIdx);
                const double* ib = A_<const double>(arena, c.oInjBasal);
                const int64_t stepNo = step0 + k + 1;
                const double t = a.dt * (double)(unsigned long long)stepNo;
                for (int e = tid; e < nInj; e += nt) {
                    const double var = LEAN ? inject_host(arena, c, e, first)
                                            : inject_varying(arena, c, e, first, t, a.dt, stepNo);
                    w.B[ii[e]] += var + ib[e];
                }
            }
            if (first && nExt > 0) {
                // external channels after the injection (HSolveActive.cpp:142-148);
                // one host entry per compartment
                work_barrier(WL);
                const int32_t* xi = A_<const int32_t>(arena, c.oExtIdx);
                const double* xg = A_<const double>(arena, c.oExtG);
                const double* xge = A_<const double>(arena, c.oExtGE);
                for (int e = tid; e < nExt; e += nt) {
                    if (!LEAN && c.nXObj > 0 && xflag[xi[e]]) continue;   // in o2dExt
                    w.D[xi[e]] += xg[e];
                    w.B[xi[e]] += xge[e];
                }
            }
            if (!LEAN && c.nXObj > 0) {
                // compartments with objects sending every tick: the whole
                // external current, after the injection
                work_barrier(WL);
                for (int i = tid; i < n; i += nt)
                    if (xflag[i]) {
                        const double2 g = external_current(arena, first, c, i, step0 + k + 1);
                        w.D[i] += g.x;
                        w.B[i] += g.y;
                    }
            }
        }
        // full barrier: it also lands every outstanding global load, so the
        // level loops start from an empty queue
        __syncthreads();
        // ---- F: forward elimination of the levels below the chain, level by
        // level.  Three descriptor
        // buffers rotate: the units of level l+2 load while level l is
        // processed, so a descriptor has two levels' time to arrive (its
        // loads stay in flight across the barriers; no register copy forces
        // them to land early).  Level bounds are read a level ahead.
        const int lastU = max(nUnits - 1, 0);
        {
            int sA = Lc > 0 ? level[0] : 0;
            int sB = Lc > 0 ? level[1] : 0;
            Desc d0 = cur, d1 = load_desc(desc, min(sB + tid, lastU)), d2;
            for (int l = 0; l < Lc; l += 3) {
                const int sC = l + 2 <= Lc ? level[l + 2] : sB;
                d2 = load_desc(desc, min(sC + tid, lastU));
                sweep_level<true, SLOWG, FD>(arena, c, desc, d0, sA, sB, w);
                work_barrier(WL);
                if (l + 1 >= Lc) break;
                const int sD = l + 3 <= Lc ? level[l + 3] : sC;
                d0 = load_desc(desc, min(sD + tid, lastU));
                sweep_level<true, SLOWG, FD>(arena, c, desc, d1, sB, sC, w);
                work_barrier(WL);
                if (l + 2 >= Lc) break;
                const int sE = l + 4 <= Lc ? level[l + 4] : sD;
                d1 = load_desc(desc, min(sE + tid, lastU));
                sweep_level<true, SLOWG, FD>(arena, c, desc, d2, sC, sD, w);
                work_barrier(WL);
                sA = sD;
                sB = sE;
            }
        }
        // Vm of this thread's first compartments for phase D, loaded before
        // the chain (mixed lean kernels): they land while thread 0 runs it,
        // so phase D stores without waiting on loads
        double vPre[kVPre];
        if constexpr (VPRE) {
#pragma unroll
            for (int u = 0; u < kVPre; ++u) vPre[u] = v.V[min(tid + u * nt, n - 1)];
        }
        // ---- the chain (thread 0): forward, the root, and back
        if (tid == 0) {
            Desc dn{};
            if (nChain > 0) dn = load_desc(sChain, 0);
            for (int k = 0; k < nChain; ++k) {
                const Desc d = dn;
                if (k + 1 < nChain) dn = load_desc(sChain, k + 1);
                forward_unit<SLOWG, FD>(arena, c, d, w);
            }
            w.B[n - 1] = w.B[n - 1] / w.D[n - 1];
            if (nChain > 0) dn = load_desc(sChain, nChain - 1);
            for (int k = nChain - 1; k >= 0; --k) {
                const Desc d = dn;
                if (k > 0) dn = load_desc(sChain, k - 1);
                backward_unit<SLOWG, FD>(arena, c, d, w);
            }
        }
        // ---- B: back-substitution of the levels below the chain (VMid
        // overwrites B), in reverse level order, with the same rotation
        {
            // level Lc-1 is [sA, eA), level Lc-2 [sB, sA)
            int eA = Lc > 0 ? level[Lc] : 0;
            int sA = Lc > 0 ? level[Lc - 1] : 0;
            int sB = Lc > 1 ? level[Lc - 2] : 0;
            Desc d0 = load_desc(desc, min(sA + tid, lastU)), d1 = load_desc(desc, min(sB + tid, lastU)), d2;
            __syncthreads();
            for (int l = Lc - 1; l >= 0; l -= 3) {
                const int sC = l >= 2 ? level[l - 2] : 0;
                d2 = load_desc(desc, min(sC + tid, lastU));
                sweep_level<false, SLOWG, FD>(arena, c, desc, d0, sA, eA, w);
                work_barrier(WL);
                if (l == 0) break;
                const int sD = l >= 3 ? level[l - 3] : 0;
                d0 = load_desc(desc, min(sD + tid, lastU));
                sweep_level<false, SLOWG, FD>(arena, c, desc, d1, sB, sA, w);
                work_barrier(WL);
                if (l == 1) break;
                const int sE = l >= 4 ? level[l - 4] : 0;
                d1 = load_desc(desc, min(sE + tid, lastU));
                sweep_level<false, SLOWG, FD>(arena, c, desc, d2, sC, sB, w);
                work_barrier(WL);
                eA = sC;
                sA = sD;
                sB = sE;
            }
        }
        // ---- D: Vm, calcium
        if (SPLIT && a.deferD) {
            // the update kernel does D over the whole batch: VMid to HBM
            if constexpr (WL)
                for (int i = tid; i < n; i += nt) gwork[n + i] = w.B[i];
        } else {
            const int64_t stepNo = step0 + k + 1;
            if (SIG == kSigCa4 && v.hasPools && c.nPoolSlots == 1) {
                // two compartments per iteration, each one's loads issued
                // together
#pragma unroll 1
                for (int i = tid; i < n; i += 2 * nt) {
                    const int j = min(i + nt, n - 1);
                    Ca4In<GT> Ra, Rb;
                    ca4_load<GT>(arena, c, v, ca4Feed, i, first, Ra);
                    ca4_load<GT>(arena, c, v, ca4Feed, j, first, Rb);
                    ca4_finish<GT>(arena, c, v, ca4Feed, i, w.B[i], first, Ra);
                    if (i + nt < n) ca4_finish<GT>(arena, c, v, ca4Feed, j, w.B[j], first, Rb);
                }
            } else if (!LEAN && v.hasPools)
                for (int i = tid; i < n; i += nt) {
                    const double vmid = w.B[i];
                    const double vnew = 2 * vmid - v.V[i];
                    v.V[i] = vnew;
                    if (uniform) calcium_row<GT>(arena, c, v, sTypes, i, vmid, vnew, first);
                    else calcium_row<GT>(arena, c, v, a.types, i, vmid, vnew, first);
                }
            else {
                int i0 = tid;
                if constexpr (VPRE) {
#pragma unroll
                    for (int u = 0; u < kVPre; ++u)
                        if (tid + u * nt < n) v.V[tid + u * nt] = 2 * w.B[tid + u * nt] - vPre[u];
                    i0 = tid + kVPre * nt;
                }
                // Vm = 2 VMid - Vm; the old Vm of four compartments loaded
                // together (clamped loads: every lane issues the same loads)
                for (int i = i0; i < n; i += 4 * nt) {
                    double vo[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) vo[u] = v.V[min(i + u * nt, n - 1)];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (i + u * nt < n) v.V[i + u * nt] = 2 * w.B[i + u * nt] - vo[u];
                }
            }
            if (nSpike > 0 || nProbe > 0) {
                __syncthreads();   // V written by other threads
                tick_outputs(a, c, v.V, stepNo, pc);
            }
        }
        // the next step's phase A overwrites D/B that other threads may still
        // read; V and the gate states are touched only by their owner thread
        work_barrier(WL);
    }
    if (tid == 0) *stepCounter = step0 + a.nsteps;
}

// The fused kernel.  Its register budget is that of four waves per SIMD
// (128 VGPRs: the three rotating descriptor buffers of the level sweeps fit
// with few spills).
template <typename GT>
__global__ void __launch_bounds__(256, 4) hsolve_step_kernel(LaunchArgs a) {
    extern __shared__ double lds[];
    __shared__ ChanType sTypes[kMaxSlots];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    if (c.workInLds) cell_steps<GT, false, true, false, 0>(a, c, lds, sTypes);
    else cell_steps<GT, false, false, false, 0>(a, c, lds, sTypes);
}

// The fused kernel for cells with junction groups of four or more members
// (the group forms inlined: only the SLOWG kernels carry their registers).
template <typename GT>
__global__ void __launch_bounds__(256) hsolve_step_slow_kernel(LaunchArgs a) {
    extern __shared__ double lds[];
    __shared__ ChanType sTypes[kMaxSlots];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    if (c.workInLds) cell_steps<GT, false, true, true, 0>(a, c, lds, sTypes);
    else cell_steps<GT, false, false, true, 0>(a, c, lds, sTypes);
}

// Cells whose four channel slots are Na, Kdr, KA and CaL (kSigCa4), with
// Ca pools and (SLOWG) junction groups of four or five members: phase A
// through the fast channel pass, every load of a compartment issued
// together, instead of one round trip per gate.
template <typename GT, bool SLOWG>
__global__ void __launch_bounds__(256) hsolve_step_ca_kernel(LaunchArgs a) {
    extern __shared__ double lds[];
    __shared__ ChanType sTypes[kMaxSlots];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    if (c.workInLds) cell_steps<GT, false, true, SLOWG, 4, false, kSigCa4>(a, c, lds, sTypes);
    else cell_steps<GT, false, false, SLOWG, 4, false, kSigCa4>(a, c, lds, sTypes);
}

// Uniform two-slot cells with many compartments per thread: phase A
// through channel_row_fast (more registers, fewer workgroups per CU, the
// channel loads of a compartment in two round trips instead of one per
// gate).
// Three waves per SIMD (at most 168 VGPRs, a small spill): with the
// 52 KB of LDS of the cells it serves (n ~ 1300-2000), three workgroups fit
// on a CU; measured 3 % faster on the ensemble than two waves without
// spills, and faster than four (128 VGPRs, heavy spills).
template <typename GT>
__global__ void __launch_bounds__(256, 3) hsolve_step_fast_kernel(LaunchArgs a) {
    extern __shared__ double lds[];
    __shared__ ChanType sTypes[kMaxSlots];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    if (c.workInLds) cell_steps<GT, false, true, false, 2>(a, c, lds, sTypes);
    else cell_steps<GT, false, false, false, 2>(a, c, lds, sTypes);
}

// The same for cells with nothing beside the solver but host inputs
// (no stimulus objects, external channel objects or Ca pools); SIG = kSigHH
// for cells whose two channel slots are the squid HH pair.
template <typename GT, int SIG>
__global__ void __launch_bounds__(256, 3) hsolve_step_lean_kernel(LaunchArgs a) {
    extern __shared__ double lds[];
    __shared__ ChanType sTypes[kMaxSlots];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    if (c.workInLds) cell_steps<GT, false, true, false, 2, true, SIG>(a, c, lds, sTypes);
    else cell_steps<GT, false, false, false, 2, true, SIG>(a, c, lds, sTypes);
}

// The solve half of a tick when the channel phase runs as its own kernel.
template <typename GT, bool SLOWG>
__global__ void __launch_bounds__(256) hsolve_solve_kernel(LaunchArgs a) {
    extern __shared__ double lds[];
    __shared__ ChanType sTypes[kMaxSlots];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    if (c.workInLds) cell_steps<GT, true, true, SLOWG, 0>(a, c, lds, sTypes);
    else cell_steps<GT, true, false, SLOWG, 0>(a, c, lds, sTypes);
}

// The same solve with up to 512 threads per cell, for large cells, whose
// levels hold hundreds of units each, spread over twice as many threads.
// (1024 threads would cap the register forms of large junction groups at
// 128 VGPRs, where they spill.)
template <typename GT, bool SLOWG>
__global__ void __launch_bounds__(512) hsolve_solve_wide_kernel(LaunchArgs a) {
    extern __shared__ double lds[];
    __shared__ ChanType sTypes[kMaxSlots];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    if (c.workInLds) cell_steps<GT, true, true, SLOWG, 0>(a, c, lds, sTypes);
    else cell_steps<GT, true, false, SLOWG, 0>(a, c, lds, sTypes);
}

// The block form of the split solve for large cells (rows in HBM, HHChannel
// and 2-D channel rows already written by the tick's channel kernels):
//   hsolve_block_fwd_kernel  one workgroup per block: its rows (and the rows
//                            its subtrees feed) and descriptors staged into
//                            LDS, its units eliminated level by level with
//                            LDS-only barriers, the rows, exported term
//                            pairs and fills written back;
//   hsolve_block_top_kernel  one workgroup per cell: the top block (the units
//                            above the blocks) in LDS with the pairs the
//                            blocks exported, forward, the root, backward;
//                            the VMid of its rows written back;
//   hsolve_block_bwd_kernel  one workgroup per block: back-substitution from
//                            the VMid of the rows its subtrees feed.
// Every unit does the arithmetic of the level schedule on the same values in
// the same order (hsolve/HSolvePassive.cpp:246-401), so the result is
// identical to it, bit for bit.
// LDS of a block: descriptors (48 B per unit, first, 16-byte aligned), D and
// B of its rows, term pairs, fills, level table.
struct BlockView {
    const BlockDev* b;
    const int32_t* rows;
    double *gD, *gB, *gT, *gF, *gHJ;
    const double* gHJ0;
    const int32_t* hjMap;   // global junction coefficient of each local one
    Work w;
    int32_t* level;
    int4* desc;
};

__device__ __forceinline__ BlockView block_view(char* arena, const CellDev& c, int blk, double* lds) {
    BlockView v;
    v.b = A_<const BlockDev>(arena, c.oBlocks) + blk;
    const BlockDev& b = *v.b;
    const int R = b.nRows + b.nExt, P = b.nT + b.nTExt;
    double* gwork = A_<double>(arena, c.oWork);
    v.gD = gwork;
    v.gB = gwork + c.n;
    v.gT = gwork + 2 * c.n;
    v.gF = v.gT + 2 * c.nTerm;
    v.rows = A_<const int32_t>(arena, b.oRows);
    v.desc = reinterpret_cast<int4*>(lds);
    double* p = lds + (size_t)6 * b.nUnits;
    v.w.D = p;
    v.w.B = p + R;
    v.w.T = p + 2 * R;
    v.w.F = v.w.T + 2 * P;
    // the junction coefficients of its groups: in LDS, the initial values
    // staged in and overwritten in place by the working ones
    v.w.HJ = v.w.F + b.nF;
    v.w.HJ0 = v.w.HJ;
    v.gHJ0 = A_<const double>(arena, c.oHJ0);
    v.gHJ = v.gF + c.nFill;
    v.hjMap = A_<const int32_t>(arena, b.oHJ);
    v.level = reinterpret_cast<int32_t*>(v.w.HJ + b.nHJ);   // kBlockForms * nLevels + 1 entries
    return v;
}

// Staging copies with eight loads in flight per thread before their
// stores (a plain strided loop waits out one memory latency per element).
constexpr int kStage = 8;

template <typename T>
__device__ __forceinline__ void stage_copy(T* __restrict__ dst, const T* __restrict__ src, int count) {
    const int nt = blockDim.x;
    for (int k0 = threadIdx.x; k0 < count; k0 += kStage * nt) {
        T v[kStage];
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
            const int k = k0 + j * nt;
            v[j] = src[k < count ? k : 0];
        }
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
            const int k = k0 + j * nt;
            if (k < count) dst[k] = v[j];
        }
    }
}

// dst[k] = src[idx[k]] (and dst2[k] = src2[idx[k]] when given)
__device__ __forceinline__ void stage_gather(double* __restrict__ dst, const double* __restrict__ src,
                                             double* __restrict__ dst2, const double* __restrict__ src2,
                                             const int32_t* __restrict__ idx, int count) {
    const int nt = blockDim.x;
    for (int k0 = threadIdx.x; k0 < count; k0 += kStage * nt) {
        int r[kStage];
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
            const int k = k0 + j * nt;
            r[j] = idx[k < count ? k : 0];
        }
        double a[kStage], b[kStage];
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
            a[j] = src[r[j]];
            b[j] = dst2 ? src2[r[j]] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
            const int k = k0 + j * nt;
            if (k < count) {
                dst[k] = a[j];
                if (dst2) dst2[k] = b[j];
            }
        }
    }
}

// A block's staging with every load in flight before any LDS store: the
// first kStageAll entries per thread of each array (the index loads, then
// the data they address, then the stores), which covers the blocks' usual
// sizes; the rest of a longer array follows in the loops of stage_gather /
// stage_copy.  Rows: D and B of the first `nR` local rows; hj0: the initial
// junction coefficients (forward) or the working ones (backward); with
// fills (backward): the fills.  Then the level table and descriptors.
constexpr int kStageAll = 4;

__device__ __forceinline__ void stage_block(char* arena, const BlockView& v, int nR, const double* hjSrc,
                                            bool fills, bool imports = false) {
    const BlockDev& b = *v.b;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int32_t* fl = A_<const int32_t>(arena, b.oFill);
    const int32_t* gLv = A_<const int32_t>(arena, b.oLevel);
    const int4* gDesc = A_<const int4>(arena, b.oDesc);
    const int nLv = kBlockForms * b.nLevels + 1, nDesc = 3 * b.nUnits, nF = fills ? b.nF : 0;
    const int last = nR > 0 ? nR - 1 : 0;
    // (imports, the top: every term pair the blocks left, by global pair)
    const int32_t* im = A_<const int32_t>(arena, b.oTImp);
    const int nI = imports ? b.nTImp : 0;
    int ri[kStageAll], hi[kStageAll], fi[kStageAll], ii[kStageAll];
#pragma unroll
    for (int j = 0; j < kStageAll; ++j) {
        const int k = tid + j * nt;
        ri[j] = v.rows[min(k, last)];
        hi[j] = b.nHJ > 0 ? v.hjMap[min(k, b.nHJ - 1)] : 0;
        fi[j] = nF > 0 ? fl[min(k, nF - 1)] : 0;
        ii[j] = nI > 0 ? im[min(k, nI - 1)] : 0;
    }
    double dv[kStageAll], bv[kStageAll], hv[kStageAll], fv[kStageAll], tv[kStageAll], uv[kStageAll];
    int4 dd[kStageAll];
    int lv[kStageAll];
#pragma unroll
    for (int j = 0; j < kStageAll; ++j) {
        const int k = tid + j * nt;
        dv[j] = v.gD[ri[j]];
        bv[j] = v.gB[ri[j]];
        hv[j] = b.nHJ > 0 ? hjSrc[hi[j]] : 0.0;
        fv[j] = nF > 0 ? v.gF[fi[j]] : 0.0;
        tv[j] = nI > 0 ? v.gT[2 * ii[j]] : 0.0;
        uv[j] = nI > 0 ? v.gT[2 * ii[j] + 1] : 0.0;
        dd[j] = gDesc[min(k, nDesc - 1)];
        lv[j] = gLv[min(k, nLv - 1)];
    }
#pragma unroll
    for (int j = 0; j < kStageAll; ++j) {
        const int k = tid + j * nt;
        if (k < nR) {
            v.w.D[k] = dv[j];
            v.w.B[k] = bv[j];
        }
        if (k < b.nHJ) v.w.HJ[k] = hv[j];
        if (k < nF) v.w.F[k] = fv[j];
        if (k < nI) {
            v.w.T[2 * k] = tv[j];
            v.w.T[2 * k + 1] = uv[j];
        }
        if (k < nDesc) v.desc[k] = dd[j];
        if (k < nLv) v.level[k] = lv[j];
    }
    const int off = kStageAll * nt;
    for (int k = off + tid; k < nI; k += nt) {
        const int g = im[k];
        v.w.T[2 * k] = v.gT[2 * g];
        v.w.T[2 * k + 1] = v.gT[2 * g + 1];
    }
    if (nR > off) stage_gather(v.w.D + off, v.gD, v.w.B + off, v.gB, v.rows + off, nR - off);
    if (b.nHJ > off) stage_gather(v.w.HJ + off, hjSrc, nullptr, nullptr, v.hjMap + off, b.nHJ - off);
    if (nF > off) stage_gather(v.w.F + off, v.gF, nullptr, nullptr, fl + off, nF - off);
    if (nDesc > off) stage_copy(v.desc + off, gDesc + off, nDesc - off);
    if (nLv > off) stage_copy(v.level + off, gLv + off, nLv - off);
}


// A block's units with their work arrays in LDS.  The groups of four or
// five members are inlined here (unlike the step kernels' out-of-line
// forms), so that their row, term and coefficient accesses stay LDS
// instructions rather than generic ones.  A level's units come sorted by
// form (two or three members, four, five); each form's run is laid on
// whole wavefronts (its lanes padded to a multiple of 64), so a wavefront
// runs one form's code rather than all three under masks.
template <bool FWD, bool SLOWG, bool FD>
__device__ __forceinline__ void block_level(const BlockView& v, int l) {
    const int32_t* lv = v.level + kBlockForms * l;
    const int s0 = lv[0], s1 = lv[1], s2 = lv[2], e = lv[3];
    const int p0 = (s1 - s0 + 63) & ~63, p1 = (s2 - s1 + 63) & ~63, p2 = (e - s2 + 63) & ~63;
    for (int q = threadIdx.x; q < p0 + p1 + p2; q += blockDim.x) {
        if (q < p0) {
            if (s0 + q < s1) {
                const Desc d = load_desc(v.desc, s0 + q);
                if (FWD) fwd_small<FD>(d, v.w);
                else bwd_small<FD>(d, v.w);
            }
        } else if (q < p0 + p1) {
            if (SLOWG && s1 + q - p0 < s2) {
                const Desc d = load_desc(v.desc, s1 + q - p0);
                if (FWD) fwd_group_desc<4, FD>(d, v.w.HJ0, v.w);
                else bwd_group_desc<4, FD>(d, v.w);
            }
        } else if (SLOWG && s2 + q - p0 - p1 < e) {
            const Desc d = load_desc(v.desc, s2 + q - p0 - p1);
            if (FWD) fwd_group_desc<5, FD>(d, v.w.HJ0, v.w);
            else bwd_group_desc<5, FD>(d, v.w);
        }
    }
}

// FD: the mixed form divides with the refined reciprocal (quot<true>), the
// FP64 form exactly.
template <bool SLOWG, bool FD>
__device__ __forceinline__ void block_forward(const LaunchArgs& a, const BlockView& v) {
    for (int l = 0; l < v.b->nLevels; ++l) {
        block_level<true, SLOWG, FD>(v, l);
        lds_barrier();
    }
}

template <bool SLOWG, bool FD>
__device__ __forceinline__ void block_backward(const LaunchArgs& a, const BlockView& v) {
    for (int l = v.b->nLevels - 1; l >= 0; --l) {
        block_level<false, SLOWG, FD>(v, l);
        lds_barrier();
    }
}

template <bool SLOWG, bool FD>
__global__ void __launch_bounds__(256) hsolve_block_fwd_kernel(LaunchArgs a, const int2* __restrict__ tiles) {
    extern __shared__ double lds[];
    const int2 t = tiles[blockIdx.x];
    const CellDev& c = a.cells[t.x];
    char* arena = a.arena;
    BlockView v = block_view(arena, c, t.y, lds);
    const BlockDev& b = *v.b;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int R = b.nRows + b.nExt;
    stage_block(arena, v, R, v.gHJ0, false);
    __syncthreads();
    block_forward<SLOWG, FD>(a, v);
    // the rows (eliminated, for the way back; fed, for the top), the term
    // pairs left for the top, the fills, the working junction coefficients
    for (int k = tid; k < R; k += nt) {
        const int r = v.rows[k];
        v.gD[r] = v.w.D[k];
        v.gB[r] = v.w.B[k];
    }
    const int32_t* tx = A_<const int32_t>(arena, b.oTExt);
    for (int j = tid; j < b.nTExt; j += nt) {
        const int g = tx[j];
        v.gT[2 * g] = v.w.T[2 * (b.nT + j)];
        v.gT[2 * g + 1] = v.w.T[2 * (b.nT + j) + 1];
    }
    const int32_t* fl = A_<const int32_t>(arena, b.oFill);
    for (int j = tid; j < b.nF; j += nt) v.gF[fl[j]] = v.w.F[j];
    for (int k = tid; k < b.nHJ; k += nt) v.gHJ[v.hjMap[k]] = v.w.HJ[k];
}

template <bool SLOWG, bool FD>
__global__ void __launch_bounds__(256) hsolve_block_bwd_kernel(LaunchArgs a, const int2* __restrict__ tiles) {
    extern __shared__ double lds[];
    const int2 t = tiles[blockIdx.x];
    const CellDev& c = a.cells[t.x];
    char* arena = a.arena;
    BlockView v = block_view(arena, c, t.y, lds);
    const BlockDev& b = *v.b;
    const int tid = threadIdx.x, nt = blockDim.x;
    // the rows the subtrees feed hold their VMid (the top solved them):
    // staged with the rest (their D is staged too, and not read)
    stage_block(arena, v, b.nRows + b.nExt, v.gHJ, true);
    __syncthreads();
    block_backward<SLOWG, FD>(a, v);
    for (int k = tid; k < b.nRows; k += nt) v.gB[v.rows[k]] = v.w.B[k];
}

// The top of each cell in `order` (its record follows the blocks'): the rows
// and every term pair it uses staged in, forward, the root row (its only fed
// row, VMid = B / D), backward; VMid of its rows written back.  As the solve
// kernel of the tick, it advances the cell's step counter.
template <bool SLOWG, bool FD>
__global__ void __launch_bounds__(256) hsolve_block_top_kernel(LaunchArgs a) {
    extern __shared__ double lds[];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    char* arena = a.arena;
    BlockView v = block_view(arena, c, c.nBlocks, lds);
    const BlockDev& b = *v.b;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int R = b.nRows + b.nExt;
    stage_block(arena, v, R, v.gHJ0, false, true);
    __syncthreads();
    block_forward<SLOWG, FD>(a, v);
    if (tid == 0) v.w.B[b.nRows] = v.w.B[b.nRows] / v.w.D[b.nRows];
    lds_barrier();
    block_backward<SLOWG, FD>(a, v);
    for (int k = tid; k < R; k += nt) v.gB[v.rows[k]] = v.w.B[k];
    if (tid == 0) *A_<int64_t>(arena, c.oSteps) += a.nsteps;
}

// Rows in HBM (split mode): the row's destination for compartment i.
__device__ __forceinline__ void row_dest(char* arena, const CellDev& c, int i, double*& D, double*& B) {
    if (c.msg) {
        const int dst = A_<const int32_t>(arena, c.oRowPos)[i];
        D = A_<double>(arena, c.oDA) + dst;
        B = A_<double>(arena, c.oBA) + dst;
    } else {
        double* gwork = A_<double>(arena, c.oWork);
        D = gwork + i;
        B = gwork + c.n + i;
    }
}

// The rest of HSolveActive::updateMatrix for compartment i's row in split
// mode: the injection, then the external current (HSolveActive.cpp:117-148).
__device__ __forceinline__ void row_inputs(const LaunchArgs& a, const CellDev& c, int i, double& D, double& B) {
    const bool has2d = c.nXObj > 0 && A_<const int32_t>(a.arena, c.oXFlag)[i];
    const int e = c.nInj > 0 ? A_<const int32_t>(a.arena, c.oInjOf)[i] : -1;
    if (e >= 0) {
        // this tick ends at step (steps done + 1); the solve kernel
        // advances the counter after it
        const int64_t stepNo = *A_<const int64_t>(a.arena, c.oSteps) + 1;
        const double t = a.dt * (double)(unsigned long long)stepNo;
        const double var = inject_varying(a.arena, c, e, a.firstOfCall != 0, t, a.dt, stepNo);
        B += var + A_<const double>(a.arena, c.oInjBasal)[e];
    }
    if (has2d) {
        const double2 g = external_current(a.arena, a.firstOfCall != 0, c, i,
                                           *A_<const int64_t>(a.arena, c.oSteps) + 1);
        D += g.x;
        B += g.y;
    } else if (a.firstOfCall) {
        const int32_t* xi = A_<const int32_t>(a.arena, c.oExtIdx);
        for (int e = 0; e < c.nExt; ++e)
            if (xi[e] == i) {
                D += A_<const double>(a.arena, c.oExtG)[e];
                B += A_<const double>(a.arena, c.oExtGE)[e];
            }
    }
}

// Phase A as a kernel of its own: advanceChannels, calculateChannelCurrents
// and the updateMatrix row (HSolveActive.cpp:218-332, 68-83, 85-116) for one
// tile of `blockDim.x` consecutive compartments of one cell per workgroup,
// one compartment per thread, over the whole batch.  The rows go to the
// cell's HBM work arrays (D at the start, B after it), where the solve
// kernel of the same tick picks them up (for cells on the message-passing
// solve, at their unit-order positions).  The thread of a compartment also
// adds its injection and then its external current (host addGkEk entries on
// the first tick of a call, or the HHChannel2D sum), in the order of
// HSolveActive::updateMatrix (HSolveActive.cpp:117-148).
template <typename GT>
__global__ void __launch_bounds__(256) hsolve_channel_kernel(LaunchArgs a, const int2* __restrict__ tiles) {
    __shared__ ChanType sTypes[kMaxSlots];
    const int2 tile = tiles[blockIdx.x];
    const CellDev& c = a.cells[tile.x];
    const View<GT> v = make_view<GT>(a.arena, c, a.tsets);
    stage_types(c, a.types, sTypes);
    __syncthreads();
    const int i = tile.y + threadIdx.x;
    if (i >= v.n) return;
    Work w{};
    int dst = i;
    if (c.msg) {
        w.D = A_<double>(a.arena, c.oDA);
        w.B = A_<double>(a.arena, c.oBA);
        dst = A_<const int32_t>(a.arena, c.oRowPos)[i];
    } else {
        double* gwork = A_<double>(a.arena, c.oWork);
        w.D = gwork;
        w.B = gwork + v.n;
    }
    if (v.uniform) channel_row<GT, true>(v, sTypes, i, a.dt, w, false, dst);
    else channel_row<GT, false>(v, a.types, i, a.dt, w, false, dst);
    row_inputs(a, c, i, w.D[dst], w.B[dst]);
}

// The channel kernel for bins of uniform cells with at most MS slots (SIG:
// their slot signature, 0 for any).
template <typename GT, int MS, int SIG = 0>
__global__ void __launch_bounds__(256) hsolve_channel_fast_kernel(LaunchArgs a, const int2* __restrict__ tiles) {
    const int2 tile = tiles[blockIdx.x];
    const CellDev& c = a.cells[tile.x];
    const View<GT> v = make_view<GT>(a.arena, c, a.tsets);
    const int i = tile.y + threadIdx.x;
    if (i >= v.n) return;
    double D, B;
    if (c.uni2d) {
        // one HHChannel2D per compartment: its external current (host
        // entries included), advanced here (ext2dInRow: the bin's external
        // kernel is not launched) or read from o2dExt, and the injection
        // entry loaded with the row, then row_inputs' additions in its order
        const int e = c.nInj > 0 ? A_<const int32_t>(a.arena, c.oInjOf)[i] : -1;
        const double2 g =
            a.ext2dInRow ? external_uni2d<std::is_same<GT, float>::value>(
                               a.arena, a.tabs2d, a.dt, a.firstOfCall != 0, c, i, v.V[i],
                               A_<const double>(a.arena, c.oPoolCa))
                         : reinterpret_cast<const double2*>(A_<const double>(a.arena, c.o2dExt))[i];
        channel_row_fast<GT, MS, false, SIG>(v, c, i, v.V[i], a.dt, D, B);
        if (e >= 0) {
            const int64_t stepNo = *A_<const int64_t>(a.arena, c.oSteps) + 1;
            const double t = a.dt * (double)(unsigned long long)stepNo;
            const double var = inject_varying(a.arena, c, e, a.firstOfCall != 0, t, a.dt, stepNo);
            B += var + A_<const double>(a.arena, c.oInjBasal)[e];
        }
        D += g.x;
        B += g.y;
    } else {
        channel_row_fast<GT, MS, false, SIG>(v, c, i, v.V[i], a.dt, D, B);
        row_inputs(a, c, i, D, B);
    }
    double *Dp, *Bp;
    row_dest(a.arena, c, i, Dp, Bp);
    *Dp = D;
    *Bp = B;
}

// ---------------------------------------------------------------------------
// Message-passing solve (split mode).  The rows of a cell sit in HBM in unit
// order; a unit reads its members' rows at positions 2u, 2u+1 (prefetched a
// level ahead with its descriptor, no dependency on the descriptor), takes
// its feeders' terms from the members' LDS slots, writes its own terms for
// its last member into that member's slot, and leaves the eliminated rows
// (and the fill A[1][2]) in place for the way back.  Backward, a unit reads
// its last member's VMid from that slot, solves its members, writes their
// VMid into their slots for their feeders and updates V.  The operations
// and their order are those of the in-place elimination (fwd_small /
// bwd_small); only the home of the values in flight changes, so LDS holds
// just the live slots and many more cells fit on a CU.
// ---------------------------------------------------------------------------
struct MsgFwd {
    Desc d;
    double D0, B0, D1, B1;
};

__device__ __forceinline__ MsgFwd load_msg_fwd(const int4* __restrict__ desc, const double* DA, const double* BA,
                                               int u) {
    MsgFwd m;
    m.d = load_desc(desc, u);
    m.D0 = DA[2 * u];
    m.B0 = BA[2 * u];
    m.D1 = DA[2 * u + 1];
    m.B1 = BA[2 * u + 1];
    return m;
}

// Every lane of a level issues the same global loads and stores (lanes past
// the level's units work on a clamped copy and write to dummy slots and
// positions), and the level loops have no per-lane branches around memory
// operations: the waits the compiler places for the prefetched buffers then
// count the outstanding operations exactly and never drain the next level's
// prefetch.  Dummy targets: LDS slot nSlots, rows 2*nUnits+1 (and +2), fill
// nUnits, VMid n.
__device__ __forceinline__ void msg_forward(const MsgFwd& m, int u, bool on, double* slot, int dummySlot,
                                            int dummyRow, double* DA, double* BA, double* FU) {
    const bool three = (m.d.a.x & 15) == 3;
    const int cnt0 = (m.d.a.x >> 4) & 3, cnt1 = (m.d.a.x >> 6) & 3;
    const double a01 = desc_double(m.d, 6), a02 = desc_double(m.d, 8), a12 = desc_double(m.d, 10);
    double D0 = m.D0, B0 = m.B0;
    if (cnt0) {
        const double* t = slot + kSlotDoubles * m.d.a.y;
        D0 -= t[0];
        B0 -= t[1];
        if (cnt0 == 2) {
            D0 -= t[2];
            B0 -= t[3];
        }
    }
    const double l1 = a01 / D0;
    const double tD1 = l1 * a01;
    const double tB1 = l1 * B0;
    // second member (three-member units; computed and discarded otherwise)
    double D1 = m.D1 - tD1, B1 = m.B1 - tB1;
    if (cnt1) {
        const double* t = slot + kSlotDoubles * m.d.a.z;
        D1 -= t[0];
        B1 -= t[1];
        if (cnt1 == 2) {
            D1 -= t[2];
            B1 -= t[3];
        }
    }
    const double l2 = a02 / D0;
    const double tD2 = l2 * a02;
    const double tB2 = l2 * B0;
    const double a12u = a12 - l1 * a02;
    const double a12l = a12 - l2 * a01;
    const double l = a12l / D1;
    double* o = slot + kSlotDoubles * (on ? m.d.a.w : dummySlot);
    o[0] = three ? tD2 : tD1;
    o[1] = three ? tB2 : tB1;
    o[2] = l * a12u;
    o[3] = l * B1;
    const int r = on ? 2 * u : dummyRow;
    DA[r] = D0;
    BA[r] = B0;
    DA[r + 1] = D1;     // a two-member unit's second position is unused
    BA[r + 1] = B1;
    FU[r >> 1] = a12u;
}

struct MsgBwd {
    Desc d;
    double D0, B0, D1, B1, F;
};

__device__ __forceinline__ MsgBwd load_msg_bwd(const int4* __restrict__ desc, const double* DA, const double* BA,
                                               const double* FU, int u) {
    MsgBwd m;
    m.d = load_desc(desc, u);
    m.D0 = DA[2 * u];
    m.B0 = BA[2 * u];
    m.D1 = DA[2 * u + 1];
    m.B1 = BA[2 * u + 1];
    m.F = FU[u];
    return m;
}

__device__ __forceinline__ void msg_backward(const MsgBwd& m, bool on, double* slot, int dummySlot, double* VMid,
                                             int dummyRow) {
    const bool three = (m.d.a.x & 15) == 3;
    // m1 == m0 for two-member units: the store of v1 is then overwritten
    const int m0 = on ? (m.d.b.x & 0xffff) : dummyRow;
    const int m1 = on ? ((m.d.b.x >> 16) & 0xffff) : dummyRow;
    const double a01 = desc_double(m.d, 6), a02 = desc_double(m.d, 8);
    const double vl = slot[kSlotDoubles * m.d.a.w];
    const double v1 = (m.B1 - vl * m.F) / m.D1;
    const double acc = three ? (m.B0 - vl * a02) - v1 * a01 : m.B0 - a01 * vl;
    const double v0 = acc / m.D0;
    const int o1 = on && three && m.d.a.z >= 0 ? m.d.a.z : dummySlot;
    const int o0 = on && m.d.a.y >= 0 ? m.d.a.y : dummySlot;
    slot[kSlotDoubles * o1] = v1;
    slot[kSlotDoubles * o0] = v0;
    VMid[m1] = v1;
    VMid[m0] = v0;
}

template <typename GT>
__global__ void __launch_bounds__(256) hsolve_msg_kernel(LaunchArgs a) {
    extern __shared__ double lds[];
    __shared__ ChanType sTypes[kMaxSlots];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    char* arena = a.arena;
    const int n = c.n;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int nLevels = c.nLevels;
    const int nUnits = c.nUnits;
    const int dummySlot = c.nSlots;       // LDS slot nSlots
    const int dummyRow = 2 * nUnits + 1;  // rows 2*nUnits+1, +2 and fill nUnits
    double* slot = lds;
    int32_t* level = reinterpret_cast<int32_t*>(lds + kSlotDoubles * (c.nSlots + 1));
    {
        const int32_t* gLevel = A_<const int32_t>(arena, c.oLevelStart);
        for (int l = tid; l <= nLevels; l += nt) level[l] = gLevel[l];
    }
    const int4* desc = A_<const int4>(arena, c.oMsgDesc);
    double* DA = A_<double>(arena, c.oDA);
    double* BA = A_<double>(arena, c.oBA);
    double* FU = A_<double>(arena, c.oFillU);
    double* V = A_<double>(arena, c.oV);
    double* VMid = A_<double>(arena, c.oVMid);
    int64_t* stepCounter = A_<int64_t>(arena, c.oSteps);
    const int64_t step0 = *stepCounter;
    ProbeClock pc = probe_clock(arena, c, step0);
    const int last = nUnits - 1;
    __syncthreads();

    for (int k = 0; k < a.nsteps; ++k) {
        if (nLevels > 0) {
            // ---- forward, two alternating prefetch buffers
            int s0 = level[0], s1 = level[1];
            MsgFwd cur = load_msg_fwd(desc, DA, BA, min(tid, last));
            MsgFwd nxt;
            for (int l = 0;; l += 2) {
                nxt = load_msg_fwd(desc, DA, BA, min(s1 + tid, last));
                const int s2 = l + 2 <= nLevels ? level[l + 2] : s1;
                msg_forward(cur, s0 + tid, s0 + tid < s1, slot, dummySlot, dummyRow, DA, BA, FU);
                for (int b = s0 + nt; b < s1; b += nt) {
                    const int u = b + tid;
                    msg_forward(load_msg_fwd(desc, DA, BA, min(u, last)), u, u < s1, slot, dummySlot, dummyRow, DA, BA, FU);
                }
                lds_barrier();
                if (l + 1 >= nLevels) break;
                cur = load_msg_fwd(desc, DA, BA, min(s2 + tid, last));
                const int s3 = l + 3 <= nLevels ? level[l + 3] : s2;
                msg_forward(nxt, s1 + tid, s1 + tid < s2, slot, dummySlot, dummyRow, DA, BA, FU);
                for (int b = s1 + nt; b < s2; b += nt) {
                    const int u = b + tid;
                    msg_forward(load_msg_fwd(desc, DA, BA, min(u, last)), u, u < s2, slot, dummySlot, dummyRow, DA, BA, FU);
                }
                lds_barrier();
                if (l + 2 >= nLevels) break;
                s0 = s2;
                s1 = s3;
            }
        }
        // ---- the root row: its feeder's terms, then VMid = B / D
        if (tid == 0) {
            const int r = 2 * nUnits;
            double D = DA[r], B = BA[r];
            if (c.rootCnt) {
                const double* t = slot + kSlotDoubles * c.rootSlot;
                D -= t[0];
                B -= t[1];
                if (c.rootCnt == 2) {
                    D -= t[2];
                    B -= t[3];
                }
            }
            const double vr = B / D;
            if (c.rootCnt) slot[kSlotDoubles * c.rootSlot] = vr;
            VMid[n - 1] = vr;
        }
        // the eliminated rows were written to HBM by other threads
        __syncthreads();
        if (nLevels > 0) {
            // ---- backward, reverse level order
            int s1 = level[nLevels], s0 = level[nLevels - 1];
            int p0 = nLevels > 1 ? level[nLevels - 2] : 0;
            MsgBwd cur = load_msg_bwd(desc, DA, BA, FU, min(s0 + tid, last));
            MsgBwd nxt;
            for (int l = nLevels - 1;; l -= 2) {
                // level l from cur, level l-1 (units p0..s0) into nxt
                nxt = load_msg_bwd(desc, DA, BA, FU, min(p0 + tid, last));
                const int pp = l >= 2 ? level[l - 2] : 0;
                msg_backward(cur, s0 + tid < s1, slot, dummySlot, VMid, n);
                for (int b = s0 + nt; b < s1; b += nt) {
                    const int u = b + tid;
                    msg_backward(load_msg_bwd(desc, DA, BA, FU, min(u, last)), u < s1, slot, dummySlot, VMid, n);
                }
                lds_barrier();
                if (l == 0) break;
                // level l-1 from nxt, level l-2 (units pp..p0) into cur
                cur = load_msg_bwd(desc, DA, BA, FU, min(pp + tid, last));
                const int ppp = l >= 3 ? level[l - 3] : 0;
                msg_backward(nxt, p0 + tid < s0, slot, dummySlot, VMid, n);
                for (int b = p0 + nt; b < s0; b += nt) {
                    const int u = b + tid;
                    msg_backward(load_msg_bwd(desc, DA, BA, FU, min(u, last)), u < s0, slot, dummySlot, VMid, n);
                }
                lds_barrier();
                if (l == 1) break;
                s1 = p0;
                s0 = pp;
                p0 = ppp;
            }
        }
        // VMid was written by other threads
        __syncthreads();
        const bool first = a.firstOfCall && k == 0;
        if (a.deferD) {
            // the update kernel does D over the whole batch
        } else if (c.hasPools) {
            const View<GT> v = make_view<GT>(arena, c, a.tsets);
            stage_types(c, a.types, sTypes);
            __syncthreads();
            for (int i = tid; i < n; i += nt) {
                const double vmid = VMid[i];
                const double vnew = 2 * vmid - V[i];
                V[i] = vnew;
                if (c.uniform) calcium_row<GT>(arena, c, v, sTypes, i, vmid, vnew, first);
                else calcium_row<GT>(arena, c, v, a.types, i, vmid, vnew, first);
            }
        } else {
            for (int i = tid; i < n; i += nt) V[i] = 2 * VMid[i] - V[i];
        }
        const int64_t stepNo = step0 + k + 1;
        if (!a.deferD && (c.nSpike > 0 || c.nProbe > 0)) {
            __syncthreads();   // V written by other threads
            tick_outputs(a, c, V, stepNo, pc);
        }
        __syncthreads();
    }
    if (tid == 0) *stepCounter = step0 + a.nsteps;
}

// Phase D of a split tick as a kernel of its own (the solve kernels left
// VMid in HBM): V = 2 VMid - V (HSolvePassive::backwardSubstitute's last
// step, hsolve/HSolvePassive.cpp:395-400) and advanceCalcium
// (HSolveActive.cpp:153-216) for one compartment per thread over the bin's
// 256-compartment tiles, so that a large cell's sweep is spread over the
// whole GPU rather than one workgroup.
template <typename GT, int SIG = 0>
__global__ void __launch_bounds__(256) hsolve_update_kernel(LaunchArgs a, const int2* __restrict__ tiles) {
    __shared__ ChanType sTypes[kMaxSlots];
    const int2 tile = tiles[blockIdx.x];
    const CellDev& c = a.cells[tile.x];
    stage_types(c, a.types, sTypes);
    __syncthreads();
    const int i = tile.y + threadIdx.x;
    if (i >= c.n) return;
    const double* VMid = c.msg ? A_<const double>(a.arena, c.oVMid) : A_<const double>(a.arena, c.oWork) + c.n;
    double* V = A_<double>(a.arena, c.oV);
    if constexpr (SIG == kSigCa4) {
        // Na/Kdr/KA/CaL cells with one pool per compartment: every load of
        // the compartment issued together (ca4_load), then calcium_row's
        // arithmetic (ca4_finish)
        if (c.hasPools && c.nPoolSlots == 1) {
            int feed = 0;
            for (int q = 0; q < 4; ++q) feed |= (sTypes[q].caTgtLocal == 0 ? 1 : 0) << q;
            const View<GT> v = make_view<GT>(a.arena, c, a.tsets);
            const double vm = VMid[i];
            Ca4In<GT> R;
            ca4_load<GT>(a.arena, c, v, feed, i, a.firstOfCall != 0, R);
            ca4_finish<GT>(a.arena, c, v, feed, i, vm, a.firstOfCall != 0, R);
            return;
        }
    }
    const double vmid = VMid[i];
    const double vnew = 2 * vmid - V[i];
    V[i] = vnew;
    if (c.hasPools) {
        const View<GT> v = make_view<GT>(a.arena, c, a.tsets);
        if (c.uniform) calcium_row<GT>(a.arena, c, v, sTypes, i, vmid, vnew, a.firstOfCall != 0);
        else calcium_row<GT>(a.arena, c, v, a.types, i, vmid, vnew, a.firstOfCall != 0);
    }
}

// The tick's outputs (spike generators, probes) once the update kernel has
// written V; the solve kernel has advanced the step counter to this tick.
__global__ void __launch_bounds__(256) hsolve_outputs_kernel(LaunchArgs a) {
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    if (c.nSpike == 0 && c.nProbe == 0) return;
    const int64_t stepNo = *A_<const int64_t>(a.arena, c.oSteps);
    ProbeClock pc = probe_clock(a.arena, c, stepNo - 1);
    tick_outputs(a, c, A_<const double>(a.arena, c.oV), stepNo, pc);
}

// HSolveActive::reinitChannels (hsolve/HSolveActiveSetup.cpp:79-160): gates at
// A/B of the current V (and pool Ca).
template <typename GT>
__global__ void __launch_bounds__(256) hsolve_reinit_kernel(LaunchArgs a) {
    __shared__ ChanType sTypes[kMaxSlots];
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    const View<GT> v = make_view<GT>(a.arena, c, a.tsets);
    stage_types(c, a.types, sTypes);
    // HSolveActive::reinit order (hsolve/HSolveActiveSetup.cpp:43-76):
    // spike generators, compartments, Ca pools, then the gates at the new V/Ca
    {
        const double* initVm = A_<const double>(a.arena, c.oInitVm);
        for (int i = threadIdx.x; i < v.n; i += blockDim.x) v.V[i] = initVm[i];
        const int PS = c.nPoolSlots * v.n;
        double* pc = A_<double>(a.arena, c.oPoolC);
        double* ica = A_<double>(a.arena, c.oPoolICa);
        const double* basal = A_<const double>(a.arena, c.oPoolBasal);
        for (int o = threadIdx.x; o < PS; o += blockDim.x) {
            pc[o] = 0.0;
            v.poolCa[o] = basal[o];
            ica[o] = 0.0;
        }
        double* last = A_<double>(a.arena, c.oSpkLast);
        const double* refr = A_<const double>(a.arena, c.oSpkRefr);
        for (int k = threadIdx.x; k < c.nSpike; k += blockDim.x) last[k] = -refr[k];
        PulseDev* pl = A_<PulseDev>(a.arena, c.oPulse);
        // PulseGen::reinit (device/PulseGen.cpp:426-433)
        for (int k = threadIdx.x; k < c.nPulse; k += blockDim.x) {
            pl[k].out = pl[k].base;
            // reinit clears input_; a source reinit after it then sends its
            // baseLevel there
            pl[k].input = pl[k].src >= 0 && !pl[k].srcBefore ? pl[pl[k].src].base : 0.0;
            pl[k].prevInput = 0.0;
            pl[k].trigTime = -1.0;
        }
        if (threadIdx.x == 0) *A_<int64_t>(a.arena, c.oSteps) = 0;
    }
    __syncthreads();
    Work w{};
    if (v.uniform)
        for (int i = threadIdx.x; i < v.n; i += blockDim.x) channel_row<GT, true>(v, sTypes, i, a.dt, w, true, i);
    else
        for (int i = threadIdx.x; i < v.n; i += blockDim.x) channel_row<GT, false>(v, a.types, i, a.dt, w, true, i);
}

// The objects sending (Gk, Ek) every tick (Leakage, HHChannel2D, SynChan)
// over a bin's tiles of kChannelTile consecutive compartments of one cell
// (as the channel kernel), one compartment per thread; the tick's step (or
// channel) kernel then adds o2dExt to the rows.
__global__ void __launch_bounds__(256) hsolve_external_kernel(LaunchArgs a, const int2* __restrict__ tiles) {
    const int2 tile = tiles[blockIdx.x];
    const CellDev& c = a.cells[tile.x];
    const int i = tile.y + threadIdx.x;
    if (c.nXObj == 0 || i >= c.n) return;
    const double* V = A_<const double>(a.arena, c.oV);
    const double* poolCa = A_<const double>(a.arena, c.oPoolCa);
    // this tick ends at step (steps done + 1); the tick's solver kernel
    // advances the counter after it
    const int64_t stepNo = *A_<const int64_t>(a.arena, c.oSteps) + 1;
    external_objects(a.arena, a.tabs2d, a.dt, a.firstOfCall != 0, a.fp32 != 0, c, i, V[i], poolCa, stepNo);
}

// HHChannel2D::vReinit of the cells in `order`, launched before the
// solver's reinit kernel.
__global__ void __launch_bounds__(256) hsolve_chan2d_reinit_kernel(LaunchArgs a) {
    const CellDev& c = a.cells[a.order[blockIdx.x]];
    if (c.n2d > 0) chan2d_reinit(a.arena, a.tabs2d, c);
}

}  // namespace hsg

// Launch tables of hsg_step_cells_noalloc (per bin: its cell order, then its
// channel-kernel tiles) live in this array instead of memory the library
// allocates.  One copy per GPU, kept until the process ends; every call on a
// GPU writes the tables it needs before its launches, and calls that overlap
// on one GPU (several engines, several host threads) race on it.
constexpr long long kScratchBytes = 1LL << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
namespace {
constexpr int kMaxScratchDevices = 64;
void* g_scratchPtr[kMaxScratchDevices] = {};
std::mutex g_scratchMu;
}  // namespace

// The device address of `scratch` on `device` (which must be the current
// device), looked up once per device with hipGetSymbolAddress: in host code
// the array's name is only a stand-in, never a device address.
extern "C" hipError_t hsg_scratch_address(int device, void** ptr, long long* bytes) {
    if (device < 0 || device >= kMaxScratchDevices || !ptr || !bytes) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratchPtr[device] = p;
    }
    *ptr = g_scratchPtr[device];
    *bytes = kScratchBytes;
    return hipSuccess;
}

// host-side launchers (called from hsolve_engine.cpp)
extern "C" hipError_t hsg_launch_step(const hsg::LaunchArgs* a, int nblocks, int threads, int fp32, int slow,
                                     int fastms, int lean, hipStream_t stream) {
    const dim3 g(nblocks), b(threads);
    const size_t l = a->ldsBytes;
    if (fastms == 4) {
        if (fp32 && slow) hipLaunchKernelGGL((hsg::hsolve_step_ca_kernel<float, true>), g, b, l, stream, *a);
        else if (fp32) hipLaunchKernelGGL((hsg::hsolve_step_ca_kernel<float, false>), g, b, l, stream, *a);
        else if (slow) hipLaunchKernelGGL((hsg::hsolve_step_ca_kernel<double, true>), g, b, l, stream, *a);
        else hipLaunchKernelGGL((hsg::hsolve_step_ca_kernel<double, false>), g, b, l, stream, *a);
    } else if (fastms == 2 && !slow && lean == 2) {
        if (fp32) hipLaunchKernelGGL((hsg::hsolve_step_lean_kernel<float, hsg::kSigHH>), g, b, l, stream, *a);
        else hipLaunchKernelGGL((hsg::hsolve_step_lean_kernel<double, hsg::kSigHH>), g, b, l, stream, *a);
    } else if (fastms == 2 && !slow && lean) {
        if (fp32) hipLaunchKernelGGL((hsg::hsolve_step_lean_kernel<float, 0>), g, b, l, stream, *a);
        else hipLaunchKernelGGL((hsg::hsolve_step_lean_kernel<double, 0>), g, b, l, stream, *a);
    } else if (fastms == 2 && !slow) {
        if (fp32) hipLaunchKernelGGL((hsg::hsolve_step_fast_kernel<float>), g, b, l, stream, *a);
        else hipLaunchKernelGGL((hsg::hsolve_step_fast_kernel<double>), g, b, l, stream, *a);
    } else if (fp32 && slow) hipLaunchKernelGGL((hsg::hsolve_step_slow_kernel<float>), g, b, l, stream, *a);
    else if (fp32) hipLaunchKernelGGL((hsg::hsolve_step_kernel<float>), g, b, l, stream, *a);
    else if (slow) hipLaunchKernelGGL((hsg::hsolve_step_slow_kernel<double>), g, b, l, stream, *a);
    else hipLaunchKernelGGL((hsg::hsolve_step_kernel<double>), g, b, l, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_solve(const hsg::LaunchArgs* a, int nblocks, int threads, int fp32, int slow,
                                      hipStream_t stream) {
    const dim3 g(nblocks), b(threads);
    if (threads > 256) {
        if (threads > 512) return hipErrorInvalidConfiguration;
        if (fp32 && slow) hipLaunchKernelGGL((hsg::hsolve_solve_wide_kernel<float, true>), g, b, a->ldsBytes, stream, *a);
        else if (fp32) hipLaunchKernelGGL((hsg::hsolve_solve_wide_kernel<float, false>), g, b, a->ldsBytes, stream, *a);
        else if (slow) hipLaunchKernelGGL((hsg::hsolve_solve_wide_kernel<double, true>), g, b, a->ldsBytes, stream, *a);
        else hipLaunchKernelGGL((hsg::hsolve_solve_wide_kernel<double, false>), g, b, a->ldsBytes, stream, *a);
        return hipGetLastError();
    }
    if (fp32 && slow) hipLaunchKernelGGL((hsg::hsolve_solve_kernel<float, true>), g, b, a->ldsBytes, stream, *a);
    else if (fp32) hipLaunchKernelGGL((hsg::hsolve_solve_kernel<float, false>), g, b, a->ldsBytes, stream, *a);
    else if (slow) hipLaunchKernelGGL((hsg::hsolve_solve_kernel<double, true>), g, b, a->ldsBytes, stream, *a);
    else hipLaunchKernelGGL((hsg::hsolve_solve_kernel<double, false>), g, b, a->ldsBytes, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_msg(const hsg::LaunchArgs* a, int nblocks, int threads, int fp32,
                                    hipStream_t stream) {
    if (fp32)
        hipLaunchKernelGGL(hsg::hsolve_msg_kernel<float>, dim3(nblocks), dim3(threads), a->ldsBytes, stream, *a);
    else
        hipLaunchKernelGGL(hsg::hsolve_msg_kernel<double>, dim3(nblocks), dim3(threads), a->ldsBytes, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_channel(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, int fp32,
                                        hipStream_t stream) {
    if (fp32)
        hipLaunchKernelGGL(hsg::hsolve_channel_kernel<float>, dim3(ntiles), dim3(hsg::kChannelTile), 0, stream,
                           *a, tiles);
    else
        hipLaunchKernelGGL(hsg::hsolve_channel_kernel<double>, dim3(ntiles), dim3(hsg::kChannelTile), 0, stream,
                           *a, tiles);
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_reinit(const hsg::LaunchArgs* a, int nblocks, int fp32,
                                       hipStream_t stream) {
    if (fp32)
        hipLaunchKernelGGL(hsg::hsolve_reinit_kernel<float>, dim3(nblocks), dim3(256), 0, stream, *a);
    else
        hipLaunchKernelGGL(hsg::hsolve_reinit_kernel<double>, dim3(nblocks), dim3(256), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_channel_fast(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, int fp32,
                                             int ms, int sig, hipStream_t stream) {
    const dim3 g(ntiles), b(hsg::kChannelTile);
    if (sig == hsg::kSigCa4 && ms == 4) {
        if (fp32) hipLaunchKernelGGL((hsg::hsolve_channel_fast_kernel<float, 4, hsg::kSigCa4>), g, b, 0, stream, *a, tiles);
        else hipLaunchKernelGGL((hsg::hsolve_channel_fast_kernel<double, 4, hsg::kSigCa4>), g, b, 0, stream, *a, tiles);
    } else if (fp32) {
        if (ms <= 2) hipLaunchKernelGGL((hsg::hsolve_channel_fast_kernel<float, 2>), g, b, 0, stream, *a, tiles);
        else if (ms <= 4) hipLaunchKernelGGL((hsg::hsolve_channel_fast_kernel<float, 4>), g, b, 0, stream, *a, tiles);
        else hipLaunchKernelGGL((hsg::hsolve_channel_fast_kernel<float, 8>), g, b, 0, stream, *a, tiles);
    } else {
        if (ms <= 2) hipLaunchKernelGGL((hsg::hsolve_channel_fast_kernel<double, 2>), g, b, 0, stream, *a, tiles);
        else if (ms <= 4) hipLaunchKernelGGL((hsg::hsolve_channel_fast_kernel<double, 4>), g, b, 0, stream, *a, tiles);
        else hipLaunchKernelGGL((hsg::hsolve_channel_fast_kernel<double, 8>), g, b, 0, stream, *a, tiles);
    }
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_external(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, hipStream_t stream) {
    hipLaunchKernelGGL(hsg::hsolve_external_kernel, dim3(ntiles), dim3(hsg::kChannelTile), 0, stream, *a, tiles);
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_chan2d_reinit(const hsg::LaunchArgs* a, int nblocks, hipStream_t stream) {
    hipLaunchKernelGGL(hsg::hsolve_chan2d_reinit_kernel, dim3(nblocks), dim3(256), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_update(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, int fp32, int sig,
                                       hipStream_t stream) {
    const dim3 g(ntiles), b(hsg::kChannelTile);
    if (sig == hsg::kSigCa4) {
        if (fp32) hipLaunchKernelGGL((hsg::hsolve_update_kernel<float, hsg::kSigCa4>), g, b, 0, stream, *a, tiles);
        else hipLaunchKernelGGL((hsg::hsolve_update_kernel<double, hsg::kSigCa4>), g, b, 0, stream, *a, tiles);
    } else if (fp32) {
        hipLaunchKernelGGL(hsg::hsolve_update_kernel<float>, g, b, 0, stream, *a, tiles);
    } else {
        hipLaunchKernelGGL(hsg::hsolve_update_kernel<double>, g, b, 0, stream, *a, tiles);
    }
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_outputs(const hsg::LaunchArgs* a, int nblocks, hipStream_t stream) {
    hipLaunchKernelGGL(hsg::hsolve_outputs_kernel, dim3(nblocks), dim3(256), 0, stream, *a);
    return hipGetLastError();
}

template <bool FD>
static void launch_block(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, int slow, int fwd, int lds,
                         hipStream_t stream) {
    const dim3 g(ntiles), b(256);
    if (fwd) {
        if (slow) hipLaunchKernelGGL((hsg::hsolve_block_fwd_kernel<true, FD>), g, b, lds, stream, *a, tiles);
        else hipLaunchKernelGGL((hsg::hsolve_block_fwd_kernel<false, FD>), g, b, lds, stream, *a, tiles);
    } else {
        if (slow) hipLaunchKernelGGL((hsg::hsolve_block_bwd_kernel<true, FD>), g, b, lds, stream, *a, tiles);
        else hipLaunchKernelGGL((hsg::hsolve_block_bwd_kernel<false, FD>), g, b, lds, stream, *a, tiles);
    }
}

extern "C" hipError_t hsg_launch_block(const hsg::LaunchArgs* a, const int2* tiles, int ntiles, int slow, int fwd,
                                      int lds, hipStream_t stream) {
    if (a->fp32) launch_block<true>(a, tiles, ntiles, slow, fwd, lds, stream);
    else launch_block<false>(a, tiles, ntiles, slow, fwd, lds, stream);
    return hipGetLastError();
}

extern "C" hipError_t hsg_launch_top(const hsg::LaunchArgs* a, int nblocks, int threads, int slow, int lds,
                                    hipStream_t stream) {
    const dim3 g(nblocks), b(threads);
    if (a->fp32) {
        if (slow) hipLaunchKernelGGL((hsg::hsolve_block_top_kernel<true, true>), g, b, lds, stream, *a);
        else hipLaunchKernelGGL((hsg::hsolve_block_top_kernel<false, true>), g, b, lds, stream, *a);
    } else {
        if (slow) hipLaunchKernelGGL((hsg::hsolve_block_top_kernel<true, false>), g, b, lds, stream, *a);
        else hipLaunchKernelGGL((hsg::hsolve_block_top_kernel<false, false>), g, b, lds, stream, *a);
    }
    return hipGetLastError();
}

extern "C" hipError_t hsg_set_max_lds(int bytes) {
    const void* kernels[] = {
        (const void*)hsg::hsolve_step_kernel<double>, (const void*)hsg::hsolve_step_kernel<float>,
        (const void*)hsg::hsolve_step_lean_kernel<double, 0>, (const void*)hsg::hsolve_step_lean_kernel<float, 0>,
        (const void*)hsg::hsolve_step_lean_kernel<double, hsg::kSigHH>,
        (const void*)hsg::hsolve_step_lean_kernel<float, hsg::kSigHH>,
        (const void*)hsg::hsolve_step_slow_kernel<double>, (const void*)hsg::hsolve_step_slow_kernel<float>,
        (const void*)hsg::hsolve_step_fast_kernel<double>, (const void*)hsg::hsolve_step_fast_kernel<float>,
        (const void*)hsg::hsolve_solve_kernel<double, false>, (const void*)hsg::hsolve_solve_kernel<float, false>,
        (const void*)hsg::hsolve_solve_kernel<double, true>, (const void*)hsg::hsolve_solve_kernel<float, true>,
        (const void*)hsg::hsolve_msg_kernel<double>, (const void*)hsg::hsolve_msg_kernel<float>,
        (const void*)hsg::hsolve_block_fwd_kernel<false, false>, (const void*)hsg::hsolve_block_fwd_kernel<true, false>,
        (const void*)hsg::hsolve_block_bwd_kernel<false, false>, (const void*)hsg::hsolve_block_bwd_kernel<true, false>,
        (const void*)hsg::hsolve_block_top_kernel<false, false>, (const void*)hsg::hsolve_block_top_kernel<true, false>,
        (const void*)hsg::hsolve_block_fwd_kernel<false, true>, (const void*)hsg::hsolve_block_fwd_kernel<true, true>,
        (const void*)hsg::hsolve_block_bwd_kernel<false, true>, (const void*)hsg::hsolve_block_bwd_kernel<true, true>,
        (const void*)hsg::hsolve_block_top_kernel<false, true>, (const void*)hsg::hsolve_block_top_kernel<true, true>,
        (const void*)hsg::hsolve_step_ca_kernel<double, false>, (const void*)hsg::hsolve_step_ca_kernel<float, false>,
        (const void*)hsg::hsolve_step_ca_kernel<double, true>, (const void*)hsg::hsolve_step_ca_kernel<float, true>};
    for (const void* k : kernels) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
