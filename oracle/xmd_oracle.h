/*
 * xmd_oracle.h -- CPU oracle for the per-step force-and-integrate loop.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of the
 * reference's arithmetic (file:line cited at every function in
 * xmd_oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it; the product path never does.
 *
 * It takes the same xmd_params / xmd_topology structs as the engine
 * (include/xmd_hip.h) and exposes the same operations, so a parity test
 * drives both sides with identical inputs.
 */
#ifndef XMD_ORACLE_H
#define XMD_ORACLE_H
#include "../include/xmd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sys oracle_sys;

/* real_kind: 0 -> real = double (the FP64 oracle), 1 -> real = float
 * (the reference as shipped, xmd/include/xmd/types/scalar.h:4). */
int oracle_create(const xmd_params *p, const xmd_topology *t, int real_kind,
                  oracle_sys **out);
void oracle_destroy(oracle_sys *s);
const char *oracle_last_error(void);

int oracle_set_state(oracle_sys *s, int replica, const double *y,
                     uint64_t rng_state, double t);
int oracle_get_state(oracle_sys *s, int replica, double *y, double *v,
                     uint64_t *rng_state, double *t);
int oracle_set_positions(oracle_sys *s, int replica, const double *r);

int oracle_rebuild_lists(oracle_sys *s);
int oracle_get_list(oracle_sys *s, int replica, int id, int32_t *i1,
                    int32_t *i2, int64_t cap, int64_t *n);
/* Force phase only; does not change state.  F global SoA [3N]. */
int oracle_eval_forces(oracle_sys *s, uint32_t mask, double *F,
                       double *V_terms);
int oracle_qa_update(oracle_sys *s);
int oracle_get_qa_state(oracle_sys *s, int replica, xmd_qa_contact *c,
                        int64_t cap, int64_t *n, int8_t *sync);
int oracle_set_qa_state(oracle_sys *s, int replica, const xmd_qa_contact *c,
                        int64_t n, const int8_t *sync);
/* nthreads == 0: deterministic serial order.  nthreads > 0: the
 * reference's OpenMP structure (thread-private F, atomic reduce), used for
 * timing. */
int oracle_step(oracle_sys *s, int64_t nsteps, int nthreads);
int oracle_get_energy(oracle_sys *s, double *V);
int oracle_get_counters(oracle_sys *s, int64_t *out5);

/* rand_gen primitives (xmd/include/xmd/random/rand_gen.h:22-58). */
uint64_t oracle_rng_next(uint64_t *state);
float oracle_rng_uniformf(uint64_t *state);
void oracle_rand_normals(uint64_t state, int64_t n, float *out,
                         uint64_t *state_out);
/* LJ attractive wall state, as xmd_get_wall_state / xmd_set_wall_state. */
int oracle_get_wall_state(oracle_sys *h, int replica, int8_t *status, double *joint, double *ref_t);
int oracle_set_wall_state(oracle_sys *h, int replica, const int8_t *status, const double *joint,
                          const double *ref_t);
/* Pure primitives for closed-form tests: out = {V, dV/dr}. */
void oracle_prim(int kind, const double *args, double *out);
// chains of the given lengths laid out as random walks (model.cpp:92-174);
// bond <= 0 takes each bond length from orig (3 per residue); box is
// written when infer_box is set; *state advances past the draws
void oracle_morph_into_saw(uint64_t *state, int n_chains, const int *chain_len, const double *orig, double bond,
                           double density, int infer_box, double *pos, double *box);

#ifdef __cplusplus
}
#endif
#endif
