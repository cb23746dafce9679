/*
 * hsolve_amd.h — C ABI of the MI355X HSolve engine.
 *
 * One engine holds a batch of HSolve instances ("cells").  Each cell is the
 * model that one moose.HSolve object takes over in the reference: its
 * compartments, HHChannels, HHGates, CaConcs, SpikeGens and Leakage channels,
 * described in MOOSE object terms (creation order, message order) so that the
 * engine can derive the Hines numbering, junctions and lookup tables the same
 * way the reference does.  All per-step work runs on the GPU; the host only
 * sets up, launches and moves fields between steps.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative HSG_E* code on
 *     failure; hsg_last_error() then returns a thread-local message.  Nothing
 *     here aborts the process (the reference prints to cerr and returns, or
 *     asserts on an unknown Id: hsolve/HSolve.cpp:301-321,411-431,
 *     hsolve/HSolveInterface.cpp:23-36).
 *   - Object indices ("compartment 3", "channel 7", "pool 0") are positions in
 *     the arrays of the hsg_cell_desc passed to hsg_add_cell, i.e. the caller's
 *     own numbering, never the Hines numbering.
 *   - Calls are made from one host thread per engine, between steps.
 *
 * Reference interfaces replaced (paths relative to the moose-core tree):
 *   HSolve Cinfo and process/reinit ........ hsolve/HSolve.cpp:37-233
 *   HSolve::setPath / setup ................. hsolve/HSolve.cpp:286-334
 *   HSolveActive::step ...................... hsolve/HSolveActive.cpp:45-66
 *   zombie field access (get/set) ........... hsolve/HSolveInterface.cpp:66-533
 *   per-step inputs addInject/addGkEk/
 *     addConc/iCa ........................... hsolve/HSolveInterface.cpp:213-259,469-480
 *   SpikeGen spikeOut (via sendSpikes) ...... biophysics/SpikeGen.cpp:202-216,
 *                                             hsolve/HSolveActive.cpp:342-347
 */
#ifndef HSOLVE_AMD_H
#define HSOLVE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define HSG_OK            0
#define HSG_E_ARG        -1   /* bad argument (null pointer, bad index, bad size) */
#define HSG_E_MODEL      -2   /* the model cannot be taken over (see message)     */
#define HSG_E_STATE      -3   /* call not valid in the engine's current state     */
#define HSG_E_DEVICE     -4   /* a HIP runtime call failed                         */
#define HSG_E_NOMEM      -5   /* device memory budget exceeded (admission)         */
#define HSG_E_UNSUPPORTED -6  /* feature outside what the GPU path implements      */

/* ---- precision forms --------------------------------------------------- */
#define HSG_FP64   0   /* everything in FP64 (the reference's arithmetic)           */
#define HSG_MIXED  1   /* rate tables and their interpolation, Gbar, Ek, gate
                          powers and channel currents in FP32; gate states and
                          their update, Hines elimination, back-substitution and
                          Vm in FP64                                               */

/* ---- model description (MOOSE object level) ---------------------------- */

/* One HHGate prototype (biophysics/HHGate.h): tables A (alpha) and B
 * (alpha+beta) of divs+1 entries over [min, max].  `id` orders gates the way
 * std::set<Id> does in HSolveActive::createLookupTables
 * (hsolve/HSolveActiveSetup.cpp:339-358): ascending object id. */
typedef struct hsg_gate_desc {
    int64_t       id;
    double        min;
    double        max;
    uint32_t      divs;
    const double* tableA;
    const double* tableB;
} hsg_gate_desc;

/* One Compartment (biophysics/CompartmentBase.cpp:191-235). */
typedef struct hsg_compartment_desc {
    double Ra, Rm, Cm, Em, initVm, Vm, inject;
} hsg_compartment_desc;

/* One HHChannel (biophysics/HHChannelBase.cpp:31-76, ChanBase.cpp:90-117).
 * Channels of a compartment appear in the order of its "channel" messages.
 * gate[0..2] index hsg_cell_desc.gates for the X, Y, Z gates (-1: none; a gate
 * is only read when its power is > 0).  caTarget: the pool fed by IkOut (the
 * first CaConc target), caDepend: the pool sending "concen" (the last one);
 * -1 when absent (hsolve/HSolveActiveSetup.cpp:254-337). */
typedef struct hsg_channel_desc {
    int32_t compartment;
    int32_t instant;
    int32_t useConcentration;
    int32_t gate[3];
    int32_t caTarget;
    int32_t caDepend;
    double  Gbar, Ek, Xpower, Ypower, Zpower, X, Y, Z, modulation;
} hsg_channel_desc;

/* One CaConc pool (biophysics/CaConcBase.cpp:58-140). */
typedef struct hsg_caconc_desc {
    double Ca, CaBasal, tau, B, ceiling, floor;
} hsg_caconc_desc;

/* One SpikeGen fed by a compartment's VmOut (biophysics/SpikeGen.cpp:127-134). */
typedef struct hsg_spikegen_desc {
    int32_t compartment;
    int32_t edgeTriggered;
    double  threshold, refractT;
} hsg_spikegen_desc;

/* A VClamp (biophysics/VClamp.cpp) on a compartment: it senses the
 * compartment's Vm (VmOut -> sensedIn, the Vm the solver sent at the end of
 * the previous tick) and sends its PID current to the compartment's
 * injectMsg on every tick (clock tick 0, before the solver's tick 6),
 * after the PulseGens.  mode 0: PID on the error; 1: derivative on Vm;
 * 2: proportional and derivative on Vm.  gain is Kp (0: Cm / dt at reinit);
 * ti = 0 becomes dt, td < 0 becomes 0, tau = 0 becomes 5 dt at reinit
 * (VClamp::reinit, :313-347).  The command input (commandIn) is the field
 * HSG_F_VCLAMP_COMMAND. */
typedef struct hsg_vclamp_desc {
    int32_t compartment;
    int32_t mode;
    double  tau, ti, td, gain;
} hsg_vclamp_desc;

/* One Leakage object on a compartment's "channel" message.  Gk is its Gk
 * field when the solver reads the model (0 until the object has been
 * reinitialised or processed): the solver folds Ek * Gk into EmByRm, the
 * right-hand side only (hsolve/HSolvePassive.cpp:138-156).  The object is
 * not taken over: on every tick (clock tick 4, before the solver's tick 6)
 * it sends (Gbar * modulation, Ek) (biophysics/Leakage.cpp:90-95), which the
 * solver adds to the compartment's external current (HSolve::addGkEk,
 * hsolve/HSolveInterface.cpp:246-252); what it sends at reinit is cleared by
 * the solver's own reinit (hsolve/HSolveActiveSetup.cpp:43-46). */
typedef struct hsg_leak_desc {
    int32_t compartment;
    double  Gk, Ek;
    double  Gbar, modulation;
} hsg_leak_desc;

/* One SynChan (biophysics/SynChan.cpp) on a compartment's "channel" message,
 * with the SimpleSynHandler whose activationOut feeds it
 * (synapse/SimpleSynHandler.cpp): n_synapses synapses, the next n_synapses
 * entries of hsg_cell_desc.synapses (synchans in array order).  Per tick
 * the handler pops the events due (time <= t) and sends sum(weight / dt)
 * (clock tick 1), then the channel advances X and Y and sends
 * (Y * norm * modulation, Ek) to the compartment's external current
 * (clock tick 4), before the solver's tick 6 (SynChan.cpp:200-227). */
typedef struct hsg_synchan_desc {
    int32_t compartment;
    int32_t n_synapses;
    double  Gbar, Ek, tau1, tau2, modulation;
} hsg_synchan_desc;

/* One synapse of a SimpleSynHandler (synapse/Synapse.cpp:73-113): a spike
 * at time t arriving through it becomes the event (t + delay, weight). */
typedef struct hsg_synapse_desc {
    double weight, delay;
} hsg_synapse_desc;

/* One 2-D interpolation table (builtins/Interpol2D.h): (xdivs+1) x (ydivs+1)
 * values, x-major (table[ix * (ydivs+1) + iy]), over [xmin, xmax] x
 * [ymin, ymax]. */
typedef struct hsg_table2d_desc {
    uint32_t      xdivs, ydivs;
    double        xmin, xmax, ymin, ymax;
    const double* table;
} hsg_table2d_desc;

/* One HHChannel2D on a compartment's "channel" message
 * (biophysics/HHChannel2D.cpp).  HSolve does not take it over: it advances
 * on its own tick before the solver's, by exponential Euler
 * (biophysics/HHChannelBase.cpp:268-274) from the Vm and [Ca] the solver sent
 * at the end of the previous tick, and feeds its compartment through addGkEk
 * (hsolve/HSolveInterface.cpp:246-252).  The engine runs it on the device in
 * the same pass.  Per gate X, Y, Z: tables A and B (indices into
 * hsg_cell_desc.tables2d, -1: no gate) and the quantity on each table axis
 * (dep: 0 Vm, 1 conc1, 2 conc2, -1 none; biophysics/HHChannel2D.cpp:257-305).
 * conc1/conc2: the pools whose concOut feeds "concen"/"concen2" (-1: none). */
typedef struct hsg_chan2d_desc {
    int32_t compartment;
    int32_t instant;
    int32_t tableA[3];
    int32_t tableB[3];
    int32_t dep[6];          /* dep[2*g], dep[2*g+1]: axes of gate g */
    int32_t conc1, conc2;
    double  Gbar, Ek, Xpower, Ypower, Zpower, X, Y, Z, modulation;
} hsg_chan2d_desc;

/* A PulseGen whose outputOut feeds a compartment's injectMsg
 * (device/PulseGen.cpp:359-433; HSolve::addInject,
 * hsolve/HSolveInterface.cpp:213-219): n_pulses (delay, width, level)
 * entries over baseLevel, evaluated on the device before each tick at the
 * tick's end time (PulseGen runs on clock tick 0, before HSolve's tick 6), in
 * description order, and sent once at reinit (baseLevel).
 *
 * trigMode 0 runs free; 1 (EXT_TRIG) and 2 (EXT_GATE) follow the input: the
 * value the host writes (HSG_F_PULSE_INPUT, PulseGen::input; reinit sets it
 * to 0), or, with input_src >= 0, the output of that PulseGen of the cell (an
 * "output" -> "input" message), delivered as the reference's messages are:
 * the same tick's output when the source comes earlier in the list, the
 * previous tick's (at the first tick, its reinit baseLevel) when later.  A
 * source must run free with pulses that cover its period (else
 * HSG_E_UNSUPPORTED); compartment may be -1 for a PulseGen that only drives
 * another's input. */
#define HSG_MAX_PULSES 4
typedef struct hsg_pulsegen_desc {
    int32_t compartment;     /* target of injectMsg, or -1 (input source only) */
    int32_t n_pulses;        /* 1..HSG_MAX_PULSES (PulseGen::count)          */
    double  baseLevel;
    double  delay[HSG_MAX_PULSES], width[HSG_MAX_PULSES], level[HSG_MAX_PULSES];
    int32_t trigMode;        /* 0 FREE_RUN, 1 EXT_TRIG, 2 EXT_GATE            */
    int32_t input_src;       /* trigMode 1-2: PulseGen of this cell driving
                                the input, or -1; ignored in trigMode 0      */
} hsg_pulsegen_desc;

/* One cell.  axial[2*k], axial[2*k+1] = (parent, child) of the k-th
 * "axial"->"raxial" message, in creation order.  seed: the compartment
 * HSolve.target resolves to (hsolve/HSolve.cpp:315-334). */
typedef struct hsg_cell_desc {
    int32_t                     n_compartments;
    const hsg_compartment_desc* compartments;
    int32_t                     n_axial;
    const int32_t*              axial;
    int32_t                     seed;
    int32_t                     n_channels;
    const hsg_channel_desc*     channels;
    int32_t                     n_gates;
    const hsg_gate_desc*        gates;
    int32_t                     n_pools;
    const hsg_caconc_desc*      pools;
    int32_t                     n_spikegens;
    const hsg_spikegen_desc*    spikegens;
    int32_t                     n_leaks;
    const hsg_leak_desc*        leaks;
    int32_t                     n_tables2d;
    const hsg_table2d_desc*     tables2d;
    int32_t                     n_chans2d;
    const hsg_chan2d_desc*      chans2d;
    int32_t                     n_pulsegens;
    const hsg_pulsegen_desc*    pulsegens;
    int32_t                     n_synchans;
    const hsg_synchan_desc*     synchans;
    const hsg_synapse_desc*     synapses;   /* sum of the synchans' n_synapses */
    int32_t                     n_vclamps;
    const hsg_vclamp_desc*      vclamps;
} hsg_cell_desc;

/* ---- engine ------------------------------------------------------------ */

typedef struct hsg_engine hsg_engine;

typedef struct hsg_options {
    int32_t precision;        /* HSG_FP64 or HSG_MIXED                              */
    int32_t device;           /* HIP device ordinal                                 */
    int64_t mem_budget;       /* bytes of device memory the engine may hold; 0 = no
                                 limit.  hsg_add_cell fails with HSG_E_NOMEM beyond */
    int32_t steps_per_launch; /* steps fused into one kernel launch (0 = default)   */
    int32_t graph_chunk;      /* launches captured per hipGraph (0 = no graph)      */
    int32_t block_threads;    /* threads per cell workgroup, 64-256 (0 = by size)   */
    int32_t kernel_mode;      /* HSG_KERNEL_*: how a tick is split into kernels     */
} hsg_options;

/* kernel_mode values */
#define HSG_KERNEL_AUTO   0   /* engine's choice per cell: SPLIT for cells of 4096+
                                 compartments, FUSED for the rest                          */
#define HSG_KERNEL_FUSED  1   /* one workgroup per cell runs whole ticks, many per launch  */
#define HSG_KERNEL_SPLIT  2   /* per tick: a batch-wide channel kernel, then the per-cell
                                 solve kernel (Hines elimination, Vm, Ca, spikes)          */

/* Creates an engine (replaces Shell::doCreate("HSolve") and the HSolve
 * constructor, hsolve/HSolve.cpp:196-199). */
int hsg_create(const hsg_options* opts, hsg_engine** out);

/* Frees every device and host buffer (hsolve/HSolve.cpp:201-215). */
int hsg_destroy(hsg_engine* e);

/* Takes over one cell: the work of HSolve::setPath/setup with the solver's
 * dt and caAdvance fields (hsolve/HSolve.cpp:286-334,409-431;
 * hsolve/HSolveActiveSetup.cpp:17-41).  dt must be > 0.  The cell starts in
 * the state the model objects held (their Vm, X/Y/Z, Ca); call hsg_reinit to
 * reset it. *cell_id receives the engine-wide cell number. */
int hsg_add_cell(hsg_engine* e, const hsg_cell_desc* desc, double dt,
                 int32_t caAdvance, int32_t* cell_id);

/* Releases a cell's slot (serving: a finished job leaves the batch). */
int hsg_remove_cell(hsg_engine* e, int32_t cell);

/* HSolve::reinit for every cell (hsolve/HSolve.cpp:229-233,
 * hsolve/HSolveActiveSetup.cpp:43-160): V=initVm, Ca=CaBasal, gates at their
 * steady state, spike generators re-armed, time reset to 0. dt is the clock
 * tick used by later steps. */
int hsg_reinit(hsg_engine* e, double dt);

/* Same, for one cell only (a job admitted into a running batch). */
int hsg_reinit_cell(hsg_engine* e, int32_t cell);

/* Advances every live cell by nsteps ticks: HSolve::process per tick
 * (hsolve/HSolve.cpp:221-227 -> hsolve/HSolveActive.cpp:45-66). */
int hsg_step(hsg_engine* e, int64_t nsteps);

/* Advances only the given cells by their own step counts (serving). */
int hsg_step_cells(hsg_engine* e, const int32_t* cells, const int64_t* nsteps,
                   int32_t n);

/* hsg_step_cells for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc);
 * it has no counterpart in the reference.  Same arguments, semantics and
 * results as hsg_step_cells, but it calls no hipMalloc, hipFree or other
 * allocating routine: its launch tables (per size bin, the cell order and
 * the channel-kernel tiles) go to a static device array of 1 MiB per GPU,
 * held until the process ends, instead of a plan buffer; it never captures
 * a graph.  Fails with HSG_E_ARG when a call's launch tables exceed that
 * area (about 120k tiles of 256 compartments, i.e. some 30 M compartments),
 * and with HSG_E_NOMEM / HSG_E_STATE when the call would have to grow the
 * engine's own memory (arena, cell or type tables): step once with an
 * allocating call after adding cells (that first call is also the first work
 * on the engine's streams, whose queues the HIP runtime sets up on first
 * use).  Calls that overlap on one GPU (several engines on several host
 * threads) race on the array. */
int hsg_step_cells_noalloc(hsg_engine* e, const int32_t* cells, const int64_t* nsteps,
                           int32_t n);

/* ---- zombie fields ----------------------------------------------------- */
/* Compartment fields (index = compartment) */
#define HSG_F_VM        1   /* getVm/setVm           HSolveInterface.cpp:66-79   */
#define HSG_F_IM        2   /* getIm (read-only)     HSolveInterface.cpp:161-182 */
#define HSG_F_INJECT    3   /* get/setInject         HSolveInterface.cpp:190-211 */
#define HSG_F_INITVM    4   /* get/setInitVm         HSolveInterface.cpp:146-158 */
#define HSG_F_CM        5   /* get/setCm             HSolveInterface.cpp:81-96   */
#define HSG_F_EM        6   /* get/setEm             HSolveInterface.cpp:98-113  */
#define HSG_F_RM        7   /* get/setRm             HSolveInterface.cpp:115-130 */
#define HSG_F_RA        8   /* get/setRa (no matrix rebuild) :132-144            */
/* HHChannel fields (index = channel) */
#define HSG_F_GBAR     20   /* HSolveInterface.cpp:287-300 */
#define HSG_F_EK       21   /* :302-314 */
#define HSG_F_GK       22   /* :316-328 */
#define HSG_F_IK       23   /* read-only :330-339 */
#define HSG_F_X        24   /* :341-367 */
#define HSG_F_Y        25   /* :369-403 */
#define HSG_F_Z        26   /* :405-443 */
#define HSG_F_MODULATION 27 /* set only if > 0: :445-451 */
#define HSG_F_INSTANT  28   /* :273-285 */
/* CaConc fields (index = pool) */
#define HSG_F_CA       40   /* :453-467 */
#define HSG_F_CABASAL  41   /* :482-495 */
#define HSG_F_CACEIL   42   /* :505-518 */
#define HSG_F_CAFLOOR  43   /* :520-533 */
/* HHChannel2D fields (index = chans2d entry; the object's own fields,
 * biophysics/ChanBase.cpp:90-117, HHChannel2D.cpp) */
#define HSG_F_2D_GBAR  60
#define HSG_F_2D_EK    61
#define HSG_F_2D_GK    62   /* read-only: Gk of the last tick */
#define HSG_F_2D_X     64
#define HSG_F_2D_Y     65
#define HSG_F_2D_Z     66
#define HSG_F_2D_MODULATION 67
/* SynChan fields (index = synchans entry; biophysics/SynChan.cpp:117-160,
 * ChanBase.cpp:90-117).  Writing tau1, tau2 or Gbar recomputes the
 * integration constants and norm at the current dt, as the setters do. */
#define HSG_F_SYN_GBAR  80
#define HSG_F_SYN_EK    81
#define HSG_F_SYN_GK    82   /* read-only: Gk of the last tick */
#define HSG_F_SYN_TAU1  83
#define HSG_F_SYN_TAU2  84
#define HSG_F_SYN_MODULATION 85
#define HSG_F_SYN_X     86   /* read-only: the rise state X */
#define HSG_F_SYN_Y     87   /* read-only: the conductance state Y */
/* Synapse fields (index = synapses entry) */
#define HSG_F_SYNAPSE_WEIGHT 90
#define HSG_F_SYNAPSE_DELAY  91
/* VClamp fields (index = vclamps entry; biophysics/VClamp.cpp:212-285) */
#define HSG_F_VCLAMP_COMMAND 100   /* commandIn: the command voltage input */
#define HSG_F_VCLAMP_CURRENT 101   /* read-only: the current last sent */
/* PulseGen fields (index = pulsegens entry; device/PulseGen.cpp:196-204) */
#define HSG_F_PULSE_OUTPUT 110     /* read-only: output_ (the level last set) */
#define HSG_F_PULSE_INPUT  111     /* input_ (a PulseGen without input_src)  */
/* Leakage fields (index = leaks entry): the per-tick send; Gk and the
 * fold into EmByRm stay as set up */
#define HSG_F_LEAK_GBAR 95
#define HSG_F_LEAK_EK   96
#define HSG_F_LEAK_MODULATION 97

/* Reads n values of `field` for objects idx[0..n-1] of `cell` into out. */
int hsg_get_field(hsg_engine* e, int32_t cell, int32_t field,
                  const int32_t* idx, int32_t n, double* out);
/* Writes n values (same indexing). */
int hsg_set_field(hsg_engine* e, int32_t cell, int32_t field,
                  const int32_t* idx, int32_t n, const double* vals);

/* ---- per-step inputs (consumed by the next step) ------------------------ */
/* HSolve::addInject: injectVarying += value (HSolveInterface.cpp:213-219). */
int hsg_add_inject(hsg_engine* e, int32_t cell, int32_t compartment, double value);
/* HSolve::addGkEk: external channel current (HSolveInterface.cpp:246-252). */
int hsg_add_gkek(hsg_engine* e, int32_t cell, int32_t compartment, double Gk, double Ek);
/* HSolve::addConc: external [Ca] seen by a channel (HSolveInterface.cpp:254-259). */
int hsg_add_conc(hsg_engine* e, int32_t cell, int32_t channel, double conc);
/* HSolve::iCa: extra Ca-pool activation (HSolveInterface.cpp:469-480). */
int hsg_add_ica(hsg_engine* e, int32_t cell, int32_t pool, double iCa);
/* Synapse::addSpike (synapse/Synapse.cpp:104-113): a spike at `time` through
 * synapse `synapse` of synchan `synchan` (both indices into the cell's
 * description) queues the event (time + delay, weight) on the synchan's
 * handler.  Events stay queued until a tick whose end time reaches them. */
int hsg_add_spike(hsg_engine* e, int32_t cell, int32_t synchan, int32_t synapse, double time);
/* A SpikeGen's spikeOut -> Synapse addSpike message between cells of the
 * engine (or within one): every spike of spikegen `spikegen` of src_cell at
 * time t = dt * step becomes an event on dst_cell's synchan at t + delay.
 * hsg_step then advances the batch in stretches no longer than the shortest
 * such connection's delay allows (at least one tick), routing the spikes of
 * each stretch before the next. */
int hsg_connect_spike(hsg_engine* e, int32_t src_cell, int32_t spikegen, int32_t dst_cell,
                      int32_t synchan, int32_t synapse);
/* PulseGen field writes (setCount/setLevel/setWidth/setDelay/setBaseLevel,
 * device/PulseGen.cpp:243-340) on the cell's pulsegen-th PulseGen, taking
 * effect from the next step; its compartment must stay the one it was added
 * with.  The current output level is kept. */
int hsg_set_pulsegen(hsg_engine* e, int32_t cell, int32_t pulsegen, const hsg_pulsegen_desc* desc);

/* ---- outputs ------------------------------------------------------------- */
typedef struct hsg_spike {
    int32_t cell;
    int32_t spikegen;   /* index into the cell's spikegens */
    int64_t step;       /* 1-based tick index since reinit; time = step * dt */
} hsg_spike;

/* Copies out (and clears) the spikes recorded since the last call.  *n_out
 * receives how many were written (at most cap); *dropped the number lost to
 * ring-buffer overflow since the last call. */
int hsg_read_spikes(hsg_engine* e, hsg_spike* out, int32_t cap, int32_t* n_out,
                    int64_t* dropped);

/* Records Vm of the listed compartments of `cell` at the end of every step
 * into a device trace of `capacity` samples each (the Table-on-tick-8
 * semantics of builtins/Table.cpp:255-285).  Replaces any earlier probe set of
 * that cell. */
int hsg_set_probes(hsg_engine* e, int32_t cell, const int32_t* compartments,
                   int32_t n, int64_t capacity);
/* As hsg_set_probes, sampling only every `every`-th step (steps counted from
 * reinit): a Table on a clock tick `every` times the solver's dt, processed
 * after the solver in the same step (Table::process, builtins/Table.cpp:254-
 * 285; tick order and rates, scheduling/Clock.cpp:764-775). */
int hsg_set_probes_every(hsg_engine* e, int32_t cell, const int32_t* compartments,
                         int32_t n, int64_t capacity, int32_t every);
/* Copies the recorded samples: out[k*n + j] = sample k of probe j. *n_samples
 * receives the number of samples (<= capacity).  Clears the trace. */
/* Probes on other fields as well, sampled on the device like Vm probes
 * (a Table on CaConc getCa, VClamp getCurrent or SynChan getGk,
 * builtins/Table.cpp:254-285): fields[k] is HSG_F_VM (idx = compartment),
 * HSG_F_CA (idx = pool), HSG_F_VCLAMP_CURRENT (idx = vclamp) or
 * HSG_F_SYN_GK (idx = synchan).  Replaces the cell's probe set;
 * hsg_read_probes returns the columns in this order. */
int hsg_set_probe_fields(hsg_engine* e, int32_t cell, const int32_t* fields, const int32_t* idx,
                         int32_t n, int64_t capacity, int32_t every);
int hsg_read_probes(hsg_engine* e, int32_t cell, double* out, int64_t cap_samples,
                    int64_t* n_samples);

/* ---- introspection ------------------------------------------------------- */
typedef struct hsg_cell_info {
    int32_t n_compartments, n_channels, n_gates, n_pools, n_spikegens;
    int32_t n_levels;         /* elimination levels of the parallel schedule */
    int32_t max_group;        /* largest junction group (members)            */
    int64_t device_bytes;     /* device memory attributed to the cell        */
    int64_t steps_done;
    /* the lookup grids setup derived from the gates, which the reference
     * writes back into the HSolve's vMin/vMax/vDiv and caMin/caMax/caDiv
     * fields (hsolve/HSolveActiveSetup.cpp:366-409) */
    double  vMin, vMax, caMin, caMax;
    int32_t vDiv, caDiv;
} hsg_cell_info;
int hsg_cell_info_get(hsg_engine* e, int32_t cell, hsg_cell_info* out);

/* Hines index of each compartment (out[i] for compartment i): the numbering of
 * HSolvePassive::walkTree (hsolve/HSolvePassive.cpp:45-98). */
int hsg_hines_order(hsg_engine* e, int32_t cell, int32_t* out);

/* The launch plan of the live set (no counterpart in the reference, which
 * steps each cell on its own): the size bins the engine launches
 * concurrently, one per stream, with their launch shape and the duration of
 * each bin's dispatches in the last launch round of the most recent
 * hsg_step (0 where that round was a graph replay).  *n_out receives the
 * number of bins (at most cap are written). */
typedef struct hsg_bin_info {
    int32_t cells;
    int32_t threads;          /* threads per cell workgroup                     */
    int32_t solve_threads;    /* split bins: threads per cell of the solve      */
    int32_t lds_bytes;        /* dynamic LDS per workgroup                      */
    int32_t split;            /* 1: channel kernel + per-cell solve every tick  */
    int32_t msg;              /* 1: message-passing solve (split)               */
    int32_t fast_slots;       /* channel pass for uniform cells of up to this
                                 many slots (0: the generic pass)                */
    int32_t slow_groups;      /* 1: cells with junction groups of 4+ members    */
    int32_t ext;              /* 1: HHChannel2D objects (external kernel/tick)  */
    int32_t round_steps;      /* ticks in the last timed round                  */
    int64_t compartments;
    double  round_ms;         /* duration of the bin's dispatches in that round */
    char    kernel[64];       /* the per-cell kernel the bin launches           */
} hsg_bin_info;
int hsg_plan_info(hsg_engine* e, hsg_bin_info* out, int32_t cap, int32_t* n_out);

/* Launch shape of bin `bin` of the current plan (its index in
 * hsg_plan_info), for a caller that has measured the alternatives (the
 * autotuner): threads per cell workgroup (64, 128 or 256; 0: the engine's
 * choice by cell size) and the fast channel pass (1 on, 0 off, -1 the
 * engine's choice; on only where the bin's cells allow it: uniform cells
 * with integer powers).  The choice stays with the bin's class (kernel form,
 * LDS class, pair half) across later plans until changed; bin = -1 clears
 * every choice.  The next step re-plans.  No counterpart in the reference,
 * which has one cell per solver and no launch shapes (hsolve/HSolve.cpp:221). */
int hsg_set_bin_shape(hsg_engine* e, int32_t bin, int32_t threads, int32_t fast_pass);

/* Blocks until all queued device work of the engine is done. */
int hsg_sync(hsg_engine* e);

/* Kernel timing of the most recent hsg_step/hsg_step_cells call, from HIP
 * events recorded on the engine's stream around each launch: number of
 * launches and their summed and mean duration in milliseconds. */
int hsg_last_step_timing(hsg_engine* e, int32_t* launches, double* total_ms,
                         double* mean_ms);

const char* hsg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HSOLVE_AMD_H */
