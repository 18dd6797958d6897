/*
 * xmd_hip.h -- C ABI of the MI355X engine for xmd's per-step
 * force-and-integrate loop.
 *
 * The reference runs this loop as OpenMP functors wired in
 * xmd/src/freeform/main.cpp:909-1077: every force term is a pair of
 * functors, update_X (rebuilds the term's list from the Verlet list) and
 * eval_X_forces (accumulates F and V), and lang_pc_step integrates.
 * This header replaces those functors with plain C entry points over
 * caller-owned host arrays.  No C++ or torch types cross the boundary.
 *
 * Conventions
 *  - Units are xmd's internal units (xmd/include/xmd/utils/units.h:9-58):
 *    length 5 A (f77unit), time tau, energy eps, k_B = 1, charge e.
 *  - Bead arrays are global over all replicas (replicas are concatenated);
 *    every index that crosses the boundary is local to its replica unless
 *    the parameter name says "global".
 *  - Vectors are SoA: x[0..n), y[0..n), z[0..n) in one buffer of 3n.
 *  - Errors: a status code plus xmd_last_error() (replaces the reference's
 *    std::runtime_error / std::out_of_range, xmd/include/xmd/store/store.h:47).
 *  - Errors found on the device (non-finite positions, a zero uniform draw
 *    for the noise, a list capacity exceeded) belong to one replica: that
 *    replica is halted -- every later step leaves its state as it is --
 *    while the other replicas go on unchanged.  The call that first finds
 *    the error returns it, naming the replica; xmd_replica_status reports
 *    it afterwards; xmd_set_state, xmd_set_positions or xmd_set_qa_state of
 *    the replica clears it.  xmd_eval_forces and xmd_qa_update refuse to
 *    run while a replica is halted.  A replica's state after the step that
 *    failed is defined only for capacity errors of the list rebuild (the
 *    state at the start of the next step, with incomplete lists).
 *  - One host thread per engine; the engine owns all device memory.  It
 *    splits its replicas into stream groups (xmd_params.stream_groups), each
 *    with its own HIP stream and captured step graphs; xmd_step runs the
 *    groups side by side and returns when all are done, with at most a few
 *    graph launches (16 steps each) queued per group at any time.
 */
#ifndef XMD_HIP_H
#define XMD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XMD_NUM_AA 20
#define XMD_NUM_SS (XMD_NUM_AA * XMD_NUM_AA)
/* contact_type values, xmd/include/xmd/forces/qa/contact_type.h:10-37 */
#define XMD_QA_NUM_TYPES (4 + XMD_NUM_SS)
/* upper bound of QA contacts one bead can hold (sum of its contact limits) */
#define XMD_QA_MAX_SLOTS 16

typedef enum xmd_status {
    XMD_OK = 0,
    XMD_ERR_ARG = 1,        /* bad argument / inconsistent topology       */
    XMD_ERR_HIP = 2,        /* HIP runtime error                          */
    XMD_ERR_CAPACITY = 3,   /* a list outgrew its device capacity         */
    XMD_ERR_NONFINITE = 4,  /* NaN/Inf found in positions                 */
    XMD_ERR_STATE = 5,      /* call not valid in the current engine state */
    XMD_ERR_NO_DEVICE = 6   /* no HIP device / kernels missing            */
} xmd_status;

/* Force terms.  Bit k of a term mask enables term k; the bits correspond one
 * to one to the reference's `enabled` flags (freeform_params sections). */
typedef enum xmd_term {
    XMD_TERM_TETHER = 0,    /* "tether forces"        src/forces/tether.cpp:21-36            */
    XMD_TERM_CHIRAL = 1,    /* "chirality"            src/forces/chiral.cpp:15-37            */
    XMD_TERM_NAT_ANG = 2,   /* "native angles"        src/forces/angle/native.cpp:13-39      */
    XMD_TERM_CND = 3,       /* "complex native dihedrals" src/forces/dihedral/complex_native.cpp:13-49 */
    XMD_TERM_SND = 4,       /* "simple native dihedrals"  src/forces/dihedral/simple_native.cpp:13-45  */
    XMD_TERM_PAULI = 5,     /* "Pauli exclusion"      src/forces/pauli.cpp:32-48             */
    XMD_TERM_GO = 6,        /* "native contacts"      src/forces/go.cpp:31-46                */
    XMD_TERM_CONST_ES = 7,  /* "const ES"             src/forces/es/const.cpp:13-32          */
    XMD_TERM_REL_ES = 8,    /* "relative ES"          src/forces/es/relative.cpp:13-32       */
    XMD_TERM_QA = 9,        /* "quasi-adiabatic potential" src/forces/qa/process_contacts.cpp:13-57 */
    XMD_TERM_PID = 10,      /* "pseudo-improper dihedral potential" src/forces/pid/eval_forces.cpp:14-141 */
    XMD_TERM_NAT_SS = 11,   /* "native ssbonds"       src/forces/native_ss.cpp:27-40         */
    XMD_TERM_SOLID_WALL = 12, /* "solid walls"        src/forces/walls/solid.cpp:19-30       */
    XMD_TERM_LJ_WALL = 13,  /* "LJ attractive walls"  src/forces/walls/lj_attr.cpp:7-49      */
    XMD_TERM_VEL_AFM = 14,  /* "velocity AFM"         src/forces/velocity_afm.cpp:12-24      */
    XMD_TERM_FORCE_AFM = 15,/* "force AFM"            src/forces/force_afm.cpp:11-14         */
    XMD_TERM_HEUR_ANG = 16, /* "heurestic angles"     src/forces/angle/heurestic.cpp:29-59   */
    XMD_TERM_HEUR_DIH = 17, /* "heurestic dihedrals"  src/forces/dihedral/heurestic.cpp:29-70 */
    XMD_NUM_TERMS = 18
} xmd_term;

/* Lists produced by a Verlet rebuild (xmd/include/xmd/nl/data.h:8-14 and the
 * per-term update_X outputs), in canonical (i1 < i2) lexicographic order. */
typedef enum xmd_list_id {
    XMD_LIST_NL_ALL = 0,        /* every pair kept by legacy_update        */
    XMD_LIST_NL_NATIVE = 1,     /* nl_data::native_contacts                 */
    XMD_LIST_NL_NON_NATIVE = 2, /* nl_data::non_native                      */
    XMD_LIST_GO = 3,            /* update_go_contacts output                */
    XMD_LIST_PAULI = 4,         /* update_pauli_pairs output                */
    XMD_LIST_ES = 5,            /* update_es_base output (q1*q2 != 0)       */
    XMD_LIST_PID = 6,           /* update_pid_bundles output                */
    XMD_LIST_QA_FREE = 7,       /* QA free pairs (non-vacant, set order)    */
    XMD_LIST_NL_NAT_SS = 8,     /* nl_data::native_ssbonds                  */
    XMD_LIST_NAT_SS = 9         /* update_nat_ssbonds output                */
} xmd_list_id;

typedef enum xmd_precision {
    XMD_PREC_FP64 = 0,   /* every force term in FP64 (the FP64 oracle path)     */
    XMD_PREC_MIXED = 1   /* FP32 pair/bonded math from FP64-differenced
                            displacements, FP64 sums, positions, Gear and lists */
} xmd_precision;

typedef enum xmd_lambda_version {
    XMD_LAMBDA_COSINE = 0,    /* xmd/include/xmd/forces/pid/lambda.h:36-41 */
    XMD_LAMBDA_ALGEBRAIC = 1  /* lambda.h:42-50                            */
} xmd_lambda_version;

/* Flat image of freeform_params (xmd/src/freeform/params.cpp:278-309) for
 * the sections on the step path, already converted to internal units. */
typedef struct xmd_params {
    /* integrator + langevin (defaults.yml "integrator", "langevin") */
    double dt;
    double gamma_factor;
    int32_t langevin_enabled;
    int32_t precision;          /* xmd_precision */
    int32_t rng_exact;          /* 1: noise bit-equal to glibc Box-Muller */
    int32_t stream_groups;      /* replica groups stepped on concurrent HIP
                                   streams: 0 = automatic (2 when there are
                                   at least 2 replicas), n = min(n, replicas);
                                   no counterpart in freeform_params */

    /* neighbour list ("neighbor list") */
    double pad_factor;
    double max_cutoff_override; /* > 0 replaces the max over term cutoffs */

    /* "tether forces" */
    int32_t tether_enabled; int32_t _pad1;
    double tether_H1, tether_H2, tether_def_length;
    /* "chirality" */
    int32_t chir_enabled; int32_t _pad2;
    double chir_e;
    /* "native angles" */
    int32_t nat_ang_enabled; int32_t _pad3;
    double nat_ang_k;
    /* "complex native dihedrals" / "simple native dihedrals" */
    int32_t cnd_enabled; int32_t snd_enabled;
    double cnd_CDA, cnd_CDB, snd_CDH;
    /* "Pauli exclusion" */
    int32_t pauli_enabled; int32_t _pad4;
    double pauli_r_excl, pauli_depth;
    /* "native contacts" */
    int32_t go_enabled; int32_t _pad5;
    double go_depth;
    /* "const ES" / "relative ES" */
    int32_t const_es_enabled; int32_t rel_es_enabled;
    double const_es_screen, const_es_permittivity;
    double rel_es_screen, rel_es_factor;
    /* "amino acid data": charge per aa_code */
    double aa_charge[XMD_NUM_AA];

    /* "lj force variants" (lj_variants.cpp:8-63) */
    double lj_bb_rmin, lj_bb_depth, lj_bs_rmin, lj_bs_depth;
    double ss_rmax[XMD_NUM_SS];   /* index 20*aa1 + aa2 */
    double ss_depth[XMD_NUM_SS];
    int32_t ss_sinking; int32_t _pad6;

    /* "pseudo-improper dihedral potential" */
    int32_t pid_enabled; int32_t pid_lambda_version;
    double pid_bbp_alpha, pid_bbp_psi0, pid_bbp_rmin, pid_bbp_depth;
    double pid_bbm_alpha, pid_bbm_psi0, pid_bbm_rmin, pid_bbm_depth;
    double pid_ss_alpha, pid_ss_psi0;

    /* "quasi-adiabatic potential" */
    int32_t qa_enabled; int32_t _pad7;
    double qa_phase_duration, qa_breaking_factor;
    double qa_min_cos_hr, qa_min_cos_hh, qa_max_cos_nr;
    /* per aa: back, side_all, side_polar, side_hydrophobic (sync_data order) */
    int8_t aa_limits[XMD_NUM_AA][4];
    int8_t aa_polarization[XMD_NUM_AA]; /* 0 none, 1 polar, 2 hydrophobic */
    int32_t force_lanes;        /* lanes per bead of the force kernel: 0 = 8,
                                   or 4, 8, 16 (xmd_autotune picks one by
                                   timing); no counterpart in freeform_params */
    /* "native ssbonds" (params.cpp:180-184) */
    int32_t nat_ss_enabled;
    double nat_ss_H1, nat_ss_equil_dist;
    /* "solid walls", "LJ attractive walls" (params.cpp:223-246; planes in
       xmd_topology) */
    int32_t solid_walls_enabled, lj_walls_enabled;
    double solid_walls_depth, solid_walls_cutoff;
    double lj_walls_min_dist, lj_walls_breaking_factor, lj_walls_phase_duration;
    /* "velocity AFM", "force AFM" (params.cpp:248-270; tips in xmd_topology) */
    int32_t vel_afm_enabled, force_afm_enabled;
    double vel_afm_H1, vel_afm_H2;
    /* "heurestic angles", "heurestic dihedrals" (params.cpp:69-124; the
       terms' tables, angle/heurestic.h, dihedral/heurestic.h), indexed by
       the pair type 3*c1 + c2 with c = 0 GLY, 1 PRO, 2 any other residue:
       angle V = sum_d a_d theta^d (d = 0..6, eps/rad^d); dihedral
       V = c0 + c1 sin + c2 cos + c3 sin^2 + c4 cos^2 + c5 sin cos (eps) */
    int32_t heur_ang_enabled, heur_dih_enabled;
    double heur_ang_coeffs[9][7];
    double heur_dih_coeffs[9][6];
    /* width of the neighbour rows (entries per bead): 0 = automatic (three
       times the longest row of the first list plus 64, at least 256), n > 0
       = n but at least the longest first row; at most 1024.  A rebuild that
       needs wider rows widens them and steps the replica that needed them
       alone back to the others (the same results as an engine created with
       the wider rows).  No counterpart in freeform_params, whose lists are
       vectors that grow (nl/legacy_update.cpp:28-59). */
    int32_t row_width;
    /* initial capacity of each replica's QA free-pair set: 0 = automatic
       (the first list's pairs, doubled, plus the contact capacity), n > 0 =
       n rounded up to 256.  A rebuild that finds more free pairs moves
       every replica's set into larger regions and steps the replica that
       needed it alone back to the others, as for the rows.  No counterpart
       in freeform_params (qa.free_pairs is a vector, qa/free_pair.cpp). */
    int32_t fp_capacity;
    /* capacity of each replica's QA contact set: 0 = automatic (16 per bead
       plus 1024, every bead's full budget), n > 0 = n rounded up to 256.
       Removed contacts leave vacancies; the set is compacted in order when
       its extent passes half the capacity (results do not depend on it); a
       set that outgrows the capacity halts the replica (XMD_ERR_CAPACITY).
       No counterpart in freeform_params (qa.contacts is a set). */
    int32_t ct_capacity;
} xmd_params;

/* System description.  All arrays are global over replicas, caller-owned and
 * copied at creation.  Chains never span replicas; beads of a chain are
 * consecutive and in sequence order. */
typedef struct xmd_topology {
    int32_t n_replicas;
    int32_t n_beads;              /* total over replicas                       */
    int32_t n_chains;
    int32_t n_native;             /* native (Go) contacts, global indices      */
    const int32_t *replica_offset;/* [n_replicas+1] first bead of each replica */
    const int32_t *chain_offset;  /* [n_chains+1] first bead of each chain     */
    const int8_t *atype;          /* [n_beads] aa_code                         */
    const double *r0;             /* [3*n_beads] SoA initial positions          */
    const double *box;            /* [3*n_replicas] cell lengths, 0 = open axis */
    const double *temperature;    /* [n_replicas] eps/kB                        */
    const uint64_t *rng_state;    /* [n_replicas] initial rand_gen state        */
    const int32_t *nat_i1;        /* [n_native] global, i1 < i2, sorted per replica */
    const int32_t *nat_i2;
    const double *nat_dist;
    /* bonded natives, indexed by the first bead of the term; NaN = absent */
    const double *tether_len;     /* [n_beads] bond i,i+1 (NaN -> def length) */
    const double *nat_theta;      /* [n_beads] angle i,i+1,i+2                 */
    const double *nat_phi;        /* [n_beads] dihedral i..i+3                 */
    const double *nat_chir;       /* [n_beads] chirality i..i+3                */
    const double *nat_chir_factor;/* [n_beads] 1/|r23_nat|^3                   */
    /* native disulfide bonds (model contacts of type NAT_SS), global,
       i1 < i2 within one replica, sorted per replica, disjoint from nat_* */
    int32_t n_ssbonds;
    const int32_t *ss_i1;
    const int32_t *ss_i2;
    /* wall planes, the same for every replica: 6 doubles each (origin x, y,
       z, unit normal x, y, z) */
    int32_t n_solid_walls;
    const double *solid_walls;
    int32_t n_lj_walls;
    const double *lj_walls;
    /* AFM tips, global bead indices: velocity tips pull bead idx towards
       orig + t v (orig NaN = the bead's initial position), force tips add a
       constant force */
    int32_t n_vel_afm;
    const int32_t *vel_afm_idx;
    const double *vel_afm_orig;   /* [3*n_vel_afm] */
    const double *vel_afm_v;      /* [3*n_vel_afm] */
    int32_t n_force_afm;
    const int32_t *force_afm_idx;
    const double *force_afm_F;    /* [3*n_force_afm] */
    /* [n_beads] or NULL: bit 0 = the model has an angle i,i+1,i+2 without a
       native value, bit 1 = a dihedral i..i+3 without one (the heuristic
       terms' lists); the beads must be consecutive in one chain.  Typed by
       the residues of beads i, i+1 (angle) and i+1, i+2 (dihedral). */
    const uint8_t *heur_bonded;
} xmd_topology;

/* QA contact record (xmd/include/xmd/forces/qa/contact.h:15-16). */
typedef struct xmd_qa_contact {
    int32_t i1, i2;       /* local indices, i1 < i2 */
    int16_t type;         /* contact_type value      */
    int8_t status;        /* 0 FORMING_OR_FORMED, 1 BREAKING */
    int8_t _pad;
    int8_t diff1[4], diff2[4];
    double ref_time;
} xmd_qa_contact;

typedef struct xmd_engine xmd_engine;

/* Engine life cycle.  Replaces the setup in main.cpp:141-907.
 * Returns XMD_ERR_NO_DEVICE when no HIP device is visible. */
xmd_status xmd_engine_create(const xmd_params *params, const xmd_topology *topo,
                             xmd_engine **out);
void xmd_engine_destroy(xmd_engine *eng);
/* Last error message of `eng`, or of the last failed create when NULL. */
const char *xmd_last_error(const xmd_engine *eng);
/* Library version string. */
const char *xmd_version(void);
/* Index of the first of n_draws rand_gen draws from `state` whose top 32
 * bits are zero and whose index mod 4 is set in slot_mask, or -1.
 * Draw 4*(N*k + i) + {0, 2} (slot_mask 5) of a replica of N beads
 * feeds log() for bead i at step k; a zero there makes the reference's
 * noise infinite (rand_gen.h:45-46), which the engine reproduces and
 * reports as XMD_ERR_NONFINITE.  Host-only. */
int64_t xmd_rng_first_zero_draw(uint64_t state, int64_t n_draws, uint32_t slot_mask);

/* Dynamic state of one replica: Gear derivatives y0..y5 (lang_pc.h:16),
 * y = [y0x,y0y,y0z,y1x,...,y5z] each block of n beads; rand_gen state;
 * time (true_t).  Setting state invalidates the Verlet list (the next step
 * or eval rebuilds it, like nl::verify's first call, verify.cpp:5-8). */
xmd_status xmd_set_state(xmd_engine *eng, int32_t replica, const double *y,
                         uint64_t rng_state, double t);
/* v = y1 / dt is returned in v (3n) when non-NULL. */
xmd_status xmd_get_state(xmd_engine *eng, int32_t replica, double *y, double *v,
                         uint64_t *rng_state, double *t);
/* Overwrite positions (y0) only; y1..y5 untouched.  Invalidates the list. */
xmd_status xmd_set_positions(xmd_engine *eng, int32_t replica, const double *r);

/* legacy_update + every enabled update_X (main.cpp:1063-1072) for all
 * replicas at the current positions. */
xmd_status xmd_rebuild_lists(xmd_engine *eng);
/* Copy list `id` of `replica` in canonical order.  *n receives the length;
 * at most cap entries are written (i1/i2 may be NULL to query the size). */
xmd_status xmd_get_list(xmd_engine *eng, int32_t replica, int32_t id,
                        int32_t *i1, int32_t *i2, int64_t cap, int64_t *n);

/* Force phase only (reset_vf + eval_X + reduce_vf) for the terms in
 * term_mask at the current positions; rebuilds the lists first if they are
 * invalid.  QA evaluates the current contact set without changing it.
 * F: [3*n_beads] SoA global; V_terms: [n_replicas*XMD_NUM_TERMS] or NULL. */
xmd_status xmd_eval_forces(xmd_engine *eng, uint32_t term_mask, double *F,
                           double *V_terms);

/* QA bookkeeping of one step at the current positions and time: sift
 * candidates, accept them in free-pair order, advance contact states
 * (qa.cpp:20-31 order).  Forces are not returned. */
xmd_status xmd_qa_update(xmd_engine *eng);
/* QA state of one replica.  contacts: non-vacant in set order; sync: [4n]. */
xmd_status xmd_get_qa_state(xmd_engine *eng, int32_t replica,
                            xmd_qa_contact *contacts, int64_t cap,
                            int64_t *n_contacts, int8_t *sync);
xmd_status xmd_set_qa_state(xmd_engine *eng, int32_t replica,
                            const xmd_qa_contact *contacts, int64_t n_contacts,
                            const int8_t *sync);

/* State of the LJ attractive walls of one replica (the reference's saved
 * "lj_attr.pairs", main.cpp:850-860): per bead b and wall w, at index
 * b * n_lj_walls + w, status (0 FREE, 1 FORMING_OR_FORMED, 2 BREAKING),
 * joint point (3 doubles) and reference time. */
xmd_status xmd_get_wall_state(xmd_engine *eng, int32_t replica, int8_t *status, double *joint, double *ref_t);
xmd_status xmd_set_wall_state(xmd_engine *eng, int32_t replica, const int8_t *status, const double *joint,
                              const double *ref_t);

/* Run nsteps of the loop body (main.cpp:1028-1076 without reporters) on
 * every replica that is not halted: forces, Langevin predictor-corrector,
 * verify, rebuild.  Returns the first error found during the call (the
 * message names every replica that failed); replicas without errors have
 * advanced nsteps either way. */
xmd_status xmd_step(xmd_engine *eng, int64_t nsteps);
/* XMD_OK, or the error that halted `replica` (xmd_last_error describes
 * it).  No counterpart in the reference, where any error ends the run. */
xmd_status xmd_replica_status(xmd_engine *eng, int32_t replica);
/* Potential energy of the last force phase per replica ([n_replicas]). */
xmd_status xmd_get_energy(xmd_engine *eng, double *V);
/* Counters: [0] steps taken, [1] list rebuilds summed over replicas,
 * [2] current NL entries (full rows), [3] NL capacity, [4] QA contacts. */
xmd_status xmd_get_counters(xmd_engine *eng, int64_t *out5);

/* rand_gen noise of lang_pc (lang_pc.cpp:15-16) for n consecutive beads of
 * one stream starting at `state`: out[3*k..3*k+2] = (normal_x2, normal) of
 * bead k; *state_out = state after 4n draws.  Runs on the device. */
xmd_status xmd_rand_normals(xmd_engine *eng, uint64_t state, int64_t n,
                            float *out, uint64_t *state_out);

/* Instrumented steps: run nsteps of the loop eagerly (no graph) with HIP
 * events recorded around every kernel launch on the engine stream.  For
 * each distinct kernel (at most cap) returns the launch count and the
 * summed duration in ms; names receives the kernel names, '\n'-separated.
 * No counterpart in the reference (its only timing is the progress bar,
 * xmd/src/io/show_progress_bar.cpp:36-51). */
xmd_status xmd_profile(xmd_engine *eng, int64_t nsteps, char *names, int64_t names_cap,
                       int64_t *counts, double *total_ms, int32_t cap, int32_t *n_kernels);

/* Stream of the first stream group (hipStream_t as void*); with one group
 * (stream_groups = 1, or one replica) every launch of the engine uses it. */
void *xmd_stream(xmd_engine *eng);

/* Per-replica observables of the stats report (report_stats.cpp:8-41,
 * stats/{gyration,asphericity,total_energy}.cpp, report_structure.cpp:81-85),
 * in internal units: out[8*r + k] = t, V, K, E, Rg, W, native contacts
 * active (|r1 - r2| < active_threshold 2^(-1/6) d_nat), native contacts.
 * V is the energy of the last force evaluation (a step's predicted
 * positions), as in the reference. */
xmd_status xmd_observables(xmd_engine *eng, double active_threshold, double *out);

/* xmd_observables without any device allocation, for callers that cannot
 * allocate device memory (a loop under HIP graph capture, a serving loop):
 * no counterpart in the reference.  The values are staged in a fixed device
 * array of the library (8 doubles for each of at most
 * XMD_OBS_NOALLOC_MAX_REPLICAS replicas; engines with more replicas get
 * XMD_ERR_ARG), one array per GPU, shared by every call on that GPU: calls
 * that overlap on one GPU (engines driven from two host threads) race.
 * Every call overwrites the rows of its replicas. */
#define XMD_OBS_NOALLOC_MAX_REPLICAS 4096
xmd_status xmd_observables_noalloc(xmd_engine *eng, double active_threshold, double *out);

/* Autotuner (no counterpart in the reference): times 16 * reps graph-captured
 * steps (after 16 untimed ones) with 4, 8 and 16 lanes per bead of the force
 * kernel, every stream group stepped side by side as xmd_step does, each
 * sample from the current state (the engine's device state is copied first,
 * restored before every sample and at the end: it needs as much free device
 * memory as the engine holds), in three interleaved rounds; keeps the
 * candidate with the lowest median for every later step (8, the default,
 * unless the lowest is more than 3 % below its median: the tiles sum a
 * bead's partial forces in different orders, so trajectories differ in
 * rounding between tiles) and returns it in *lpb_out and the three medians
 * (ms per step) in ms_out[3] (either may be NULL).  A sample in which a
 * replica halts counts as infinitely slow.  The simulation state is not
 * changed. */
xmd_status xmd_autotune(xmd_engine *eng, int32_t reps, int32_t *lpb_out, double *ms_out);
#ifdef __cplusplus
}
#endif
#endif
