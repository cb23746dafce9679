// ref_harness.cpp — drives the reference HSolve arithmetic, compiled from the
// reference's own sources (hsolve/HinesMatrix.cpp, HSolvePassive.cpp,
// HSolveActive.cpp, HSolveActiveSetup.cpp, RateLookup.cpp, HSolveStruct.cpp),
// to check the oracle.
//
// TEST INFRASTRUCTURE ONLY.  The MOOSE object system is not linked: the
// harness fills the solver's arrays directly (they are protected members of
// HSolveActive / HSolvePassive / HinesMatrix, so a derived class may set them),
// builds the Hines matrix with the public HinesMatrix::setup(tree, dt), and
// then calls the reference's own HSolveActive::reinit and ::step.  The arrays
// come from the oracle's flat dump (or_dump_flat), i.e. from the oracle's
// walkTree/readHHChannels/readCalcium/createLookupTables; this pins the
// per-step path and the matrix construction.
//
// usage: hsolve_ref FLAT.bin NSTEPS DT OUT.bin [REINIT(1)]
// OUT.bin: int64 n, int64 nsteps, then V (n doubles) after every step,
//          then VMid (n), state (nState), ca (nPool), Gk (nChan).
// usage: hsolve_ref --time NSTEPS DT FLAT.bin...
//        the CPU baseline: every cell reinit, then NSTEPS steps of each,
//        nothing recorded; prints the compartments and the seconds the
//        step loops took.
#include "hsolve/HSolveActive.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

namespace {

struct Reader {
    FILE* f;
    long i() { long x = 0; if (fread(&x, sizeof(long), 1, f) != 1) die(); return x; }
    double d() { double x = 0; if (fread(&x, sizeof(double), 1, f) != 1) die(); return x; }
    std::vector<double> dv() {
        long n = i();
        std::vector<double> v((size_t)n);
        if (n && fread(v.data(), sizeof(double), (size_t)n, f) != (size_t)n) die();
        return v;
    }
    std::vector<long> iv() {
        long n = i();
        std::vector<long> v((size_t)n);
        for (long k = 0; k < n; ++k) v[(size_t)k] = i();
        return v;
    }
    [[noreturn]] static void die() { fprintf(stderr, "hsolve_ref: truncated input\n"); exit(2); }
};

class Harness : public HSolveActive {
public:
    void load(Reader& r) {
        if (r.i() != 0x48534f4cL) Reader::die();
        long n = r.i();
        double dt = r.d();
        caAdvance_ = (int)r.i();
        std::vector<long> cstart = r.iv();
        std::vector<long> children = r.iv();
        std::vector<double> Ra = r.dv(), Rm = r.dv(), Cm = r.dv(), Em = r.dv(), initVm = r.dv();
        std::vector<double> V = r.dv(), CmByDt = r.dv(), EmByRm = r.dv();
        std::vector<long> hasInject = r.iv();
        std::vector<double> injBasal = r.dv(), injVar = r.dv();

        tree_.resize((size_t)n);
        for (long i = 0; i < n; ++i) {
            TreeNodeStruct& t = tree_[(size_t)i];
            for (long k = cstart[(size_t)i]; k < cstart[(size_t)i + 1]; ++k)
                t.children.push_back((unsigned int)children[(size_t)k]);
            t.Ra = Ra[(size_t)i]; t.Rm = Rm[(size_t)i]; t.Cm = Cm[(size_t)i];
            t.Em = Em[(size_t)i]; t.initVm = initVm[(size_t)i];
        }
        HinesMatrix::setup(tree_, dt);
        V_ = V;
        compartment_.resize((size_t)n);
        for (long i = 0; i < n; ++i) {
            compartment_[(size_t)i].CmByDt = CmByDt[(size_t)i];
            compartment_[(size_t)i].EmByRm = EmByRm[(size_t)i];
            if (hasInject[(size_t)i]) {
                inject_[(unsigned int)i].injectBasal = injBasal[(size_t)i];
                inject_[(unsigned int)i].injectVarying = injVar[(size_t)i];
            }
        }

        long nc = r.i();
        std::vector<double> Gbar = r.dv(), Xp = r.dv(), Yp = r.dv(), Zp = r.dv();
        std::vector<long> instant = r.iv();
        std::vector<double> modulation = r.dv(), Ek = r.dv();
        state_ = r.dv();
        std::vector<long> c2s = r.iv(), c2c = r.iv(), chCount = r.iv();
        channel_.resize((size_t)nc);
        current_.resize((size_t)nc);
        for (long c = 0; c < nc; ++c) {
            ChannelStruct& ch = channel_[(size_t)c];
            ch.Gbar_ = Gbar[(size_t)c];
            ch.setPowers(Xp[(size_t)c], Yp[(size_t)c], Zp[(size_t)c]);
            ch.instant_ = (int)instant[(size_t)c];
            ch.modulation_ = modulation[(size_t)c];
            current_[(size_t)c].Ek = Ek[(size_t)c];
            current_[(size_t)c].Gk = 0.0;
            chan2state_.push_back((unsigned int)c2s[(size_t)c]);
            chan2compt_.push_back((unsigned int)c2c[(size_t)c]);
        }
        long cum = 0;
        currentBoundary_.resize((size_t)n);
        for (long i = 0; i < n; ++i) {
            channelCount_.push_back((int)chCount[(size_t)i]);
            cum += chCount[(size_t)i];
            currentBoundary_[(size_t)i] = current_.begin() + cum;
        }

        long np = r.i();
        caConc_.resize((size_t)np);
        for (long p = 0; p < np; ++p) {
            CaConcStruct& cc = caConc_[(size_t)p];
            cc.c_ = r.d(); cc.CaBasal_ = r.d(); cc.factor1_ = r.d(); cc.factor2_ = r.d();
            cc.ceiling_ = r.d(); cc.floor_ = r.d();
        }
        ca_ = r.dv();
        caActivation_.assign((size_t)np, 0.0);
        std::vector<long> caTarget = r.iv();
        caTarget_.resize((size_t)nc);
        for (long c = 0; c < nc; ++c)
            caTarget_[(size_t)c] = caTarget[(size_t)c] < 0 ? 0 : &caActivation_[(size_t)caTarget[(size_t)c]];
        std::vector<long> caCount = r.iv();
        unsigned int maxN = 0;
        for (long i = 0; i < n; ++i) {
            caCount_.push_back((unsigned int)caCount[(size_t)i]);
            if ((unsigned int)caCount[(size_t)i] > maxN) maxN = (unsigned int)caCount[(size_t)i];
        }
        std::vector<long> cols = r.iv();
        for (long col : cols) { LookupColumn lc; lc.column = (unsigned int)col; column_.push_back(lc); }
        std::vector<long> caRow = r.iv();
        caRowCompt_.resize(maxN);
        for (long idx : caRow) caRow_.push_back(idx < 0 ? 0 : &caRowCompt_[(size_t)idx]);
        externalCalcium_ = r.dv();
        for (int t = 0; t < 2; ++t) {
            double mn = r.d(), mx = r.d();
            long nDivs = r.i(), nSpecies = r.i();
            std::vector<double> tab = r.dv();
            LookupTable lt(mn, mx, (unsigned int)nDivs, (unsigned int)nSpecies);
            long nCol = 2 * nSpecies, nPts = nDivs + 2;
            for (long s = 0; s < nSpecies; ++s) {
                std::vector<double> A((size_t)(nPts - 1)), B((size_t)(nPts - 1));
                for (long k = 0; k < nPts - 1; ++k) {
                    A[(size_t)k] = tab[(size_t)(k * nCol + 2 * s)];
                    B[(size_t)k] = tab[(size_t)(k * nCol + 2 * s + 1)];
                }
                lt.addColumns((int)s, A, B);
            }
            if (t == 0) vTable_ = lt; else caTable_ = lt;
        }
        externalCurrent_.assign(2 * (size_t)n, 0.0);
        prevExtCurr_.assign(2 * (size_t)n, 0.0);
    }

    void run(long nsteps, double dt, bool doReinit, FILE* out) {
        ProcInfo p;
        p.dt = dt;
        p.currTime = 0.0;
        if (doReinit) reinit(&p);
        long n = (long)V_.size();
        fwrite(&n, sizeof(long), 1, out);
        fwrite(&nsteps, sizeof(long), 1, out);
        for (long k = 0; k < nsteps; ++k) {
            p.currTime = dt * (double)(unsigned long)(k + 1);
            step(&p);
            fwrite(V_.data(), sizeof(double), V_.size(), out);
        }
        fwrite(VMid_.data(), sizeof(double), VMid_.size(), out);
        fwrite(state_.data(), sizeof(double), state_.size(), out);
        fwrite(ca_.data(), sizeof(double), ca_.size(), out);
        for (const CurrentStruct& c : current_) fwrite(&c.Gk, sizeof(double), 1, out);
    }

    long compartments() const { return (long)V_.size(); }

    void start(double dt) {
        ProcInfo p;
        p.dt = dt;
        p.currTime = 0.0;
        reinit(&p);
    }

    void steps(long nsteps, double dt) {
        ProcInfo p;
        p.dt = dt;
        for (long k = 0; k < nsteps; ++k) {
            p.currTime = dt * (double)(unsigned long)(k + 1);
            step(&p);
        }
    }
};

int time_cells(int argc, char** argv) {
    const long nsteps = atol(argv[2]);
    const double dt = atof(argv[3]);
    std::vector<std::unique_ptr<Harness>> cells;
    long n = 0;
    for (int a = 4; a < argc; ++a) {
        FILE* f = fopen(argv[a], "rb");
        if (!f) { perror(argv[a]); return 2; }
        Reader r{f};
        cells.push_back(std::make_unique<Harness>());
        cells.back()->load(r);
        fclose(f);
        n += cells.back()->compartments();
    }
    for (auto& h : cells) h->start(dt);
    const auto t0 = std::chrono::steady_clock::now();
    for (auto& h : cells) h->steps(nsteps, dt);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    printf("%ld %.9f\n", n, s);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 5 && strcmp(argv[1], "--time") == 0) return time_cells(argc, argv);
    if (argc < 5) {
        fprintf(stderr, "usage: %s FLAT.bin NSTEPS DT OUT.bin [REINIT]\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    Reader r{f};
    Harness h;
    h.load(r);
    fclose(f);
    long nsteps = atol(argv[2]);
    double dt = atof(argv[3]);
    bool doReinit = argc < 6 || atoi(argv[5]) != 0;
    FILE* out = fopen(argv[4], "wb");
    if (!out) { perror(argv[4]); return 2; }
    h.run(nsteps, dt, doReinit, out);
    fclose(out);
    return 0;
}
