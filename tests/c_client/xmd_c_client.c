/*
 * xmd_c_client.c -- a plain C caller of the engine's C ABI
 * (include/xmd_hip.h), the way xmd's own driver would call it
 * (xmd/src/freeform/main.cpp:1009-1076): fill xmd_params and xmd_topology,
 * create the engine, step, read the state back, destroy.  No Python and no
 * torch: the test beside it (test_gpu_c_client.py) checks that this caller
 * gets the same bits as the ctypes path and the oracle's trajectory.
 *
 *   xmd_c_client run <out.bin>   one replica of two 40-bead helices (tether,
 *                                chirality, Go contacts, Pauli, Langevin),
 *                                200 steps; writes the inputs and the results
 *   xmd_c_client errors          malformed inputs are refused with
 *                                XMD_ERR_ARG and a message
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xmd_hip.h"

#define NCH 2
#define LEN 40
#define N (NCH * LEN)
#define ANG (1.0 / 5.0) /* one angstrom in the internal length unit */

static double r0[3 * N], nat_dist[N * N], tether[N], theta[N], phi[N], chir[N], chir_f[N];
static int32_t nat_i1[N * N], nat_i2[N * N];
static int8_t atype[N];
static int32_t rep_off[2] = {0, N}, chain_off[NCH + 1] = {0, LEN, 2 * LEN};

static double P3(int d, int i) { return r0[d * N + i]; }

/* Two right-handed helices (radius 2.3 A, 100 degrees and 1.5 A per bead:
   consecutive beads 3.83 A apart) 20 A apart; natives: pairs at least 3
   apart in sequence closer than 7.5 A; chirality of every quad. */
static int build_system(void) {
    for (int c = 0; c < NCH; ++c)
        for (int k = 0; k < LEN; ++k) {
            int i = c * LEN + k;
            double a = k * 100.0 * M_PI / 180.0;
            r0[0 * N + i] = (2.3 * cos(a) + 20.0 * c) * ANG;
            r0[1 * N + i] = 2.3 * sin(a) * ANG;
            r0[2 * N + i] = 1.5 * k * ANG;
            atype[i] = (int8_t)((7 * i + 3) % XMD_NUM_AA);
            tether[i] = theta[i] = phi[i] = chir[i] = chir_f[i] = NAN;
        }
    int nn = 0;
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) {
            if (i / LEN == j / LEN && j - i < 3) continue;
            double dx = P3(0, j) - P3(0, i), dy = P3(1, j) - P3(1, i), dz = P3(2, j) - P3(2, i);
            double d = sqrt(dx * dx + dy * dy + dz * dz);
            if (d < 7.5 * ANG) {
                nat_i1[nn] = i;
                nat_i2[nn] = j;
                nat_dist[nn] = d;
                ++nn;
            }
        }
    for (int c = 0; c < NCH; ++c)
        for (int k = 0; k + 3 < LEN; ++k) {
            int i = c * LEN + k;
            double r12[3], r23[3], r34[3];
            for (int d = 0; d < 3; ++d) {
                r12[d] = P3(d, i + 1) - P3(d, i);
                r23[d] = P3(d, i + 2) - P3(d, i + 1);
                r34[d] = P3(d, i + 3) - P3(d, i + 2);
            }
            double x[3] = {r23[1] * r34[2] - r23[2] * r34[1], r23[2] * r34[0] - r23[0] * r34[2],
                           r23[0] * r34[1] - r23[1] * r34[0]};
            double n23 = sqrt(r23[0] * r23[0] + r23[1] * r23[1] + r23[2] * r23[2]);
            chir_f[i] = 1.0 / (n23 * n23 * n23);
            chir[i] = (r12[0] * x[0] + r12[1] * x[1] + r12[2] * x[2]) * chir_f[i];
        }
    return nn;
}

static void set_params(xmd_params *p) {
    memset(p, 0, sizeof *p);
    p->dt = 0.005;
    p->gamma_factor = 2.0;
    p->langevin_enabled = 1;
    p->precision = XMD_PREC_FP64;
    p->rng_exact = 1;
    p->pad_factor = 0.25;
    p->tether_enabled = 1;
    p->tether_H1 = 100.0 / (ANG * ANG); /* 100 eps/A^2 */
    p->tether_def_length = 3.8 * ANG;
    p->chir_enabled = 1;
    p->chir_e = 1.0;
    p->pauli_enabled = 1;
    p->pauli_r_excl = 4.0 * ANG;
    p->pauli_depth = 1.0;
    p->go_enabled = 1;
    p->go_depth = 1.0;
}

static void set_topology(xmd_topology *t, int nn, const double *box, const double *temp, const uint64_t *seed) {
    memset(t, 0, sizeof *t);
    t->n_replicas = 1;
    t->n_beads = N;
    t->n_chains = NCH;
    t->n_native = nn;
    t->replica_offset = rep_off;
    t->chain_offset = chain_off;
    t->atype = atype;
    t->r0 = r0;
    t->box = box;
    t->temperature = temp;
    t->rng_state = seed;
    t->nat_i1 = nat_i1;
    t->nat_i2 = nat_i2;
    t->nat_dist = nat_dist;
    t->tether_len = tether;
    t->nat_theta = theta;
    t->nat_phi = phi;
    t->nat_chir = chir;
    t->nat_chir_factor = chir_f;
}

#define CHECK(call)                                                                                 \
    do {                                                                                            \
        xmd_status st_ = (call);                                                                    \
        if (st_ != XMD_OK) {                                                                        \
            fprintf(stderr, "%s failed: %d %s\n", #call, (int)st_, xmd_last_error(eng));            \
            return 1;                                                                               \
        }                                                                                           \
    } while (0)

static int run(const char *path) {
    const double box[3] = {0, 0, 0}, temp = 0.35;
    const uint64_t seed = 0x5EED00C1ull;
    const int nn = build_system();
    xmd_params p;
    xmd_topology t;
    set_params(&p);
    set_topology(&t, nn, box, &temp, &seed);
    xmd_engine *eng = NULL;
    if (xmd_engine_create(&p, &t, &eng) != XMD_OK) {
        fprintf(stderr, "create failed: %s\n", xmd_last_error(NULL));
        return 1;
    }
    static double F[3 * N], Vt[XMD_NUM_TERMS], y[18 * N];
    CHECK(xmd_eval_forces(eng, 0xffffffffu, F, Vt));
    CHECK(xmd_step(eng, 200));
    uint64_t rng = 0;
    double tnow = 0, V = 0;
    CHECK(xmd_get_state(eng, 0, y, NULL, &rng, &tnow));
    CHECK(xmd_get_energy(eng, &V));
    int64_t cnt[5];
    CHECK(xmd_get_counters(eng, cnt));
    xmd_engine_destroy(eng);

    FILE *f = fopen(path, "wb");
    if (!f) return 1;
    int32_t hdr[4] = {N, NCH, nn, XMD_NUM_TERMS};
    fwrite(hdr, sizeof hdr, 1, f);
    fwrite(&p, sizeof p, 1, f);
    fwrite(&seed, sizeof seed, 1, f);
    fwrite(&temp, sizeof temp, 1, f);
    fwrite(r0, sizeof(double), 3 * N, f);
    fwrite(atype, 1, N, f);
    fwrite(nat_i1, sizeof(int32_t), nn, f);
    fwrite(nat_i2, sizeof(int32_t), nn, f);
    fwrite(nat_dist, sizeof(double), nn, f);
    fwrite(chir, sizeof(double), N, f);
    fwrite(chir_f, sizeof(double), N, f);
    fwrite(F, sizeof(double), 3 * N, f);
    fwrite(Vt, sizeof(double), XMD_NUM_TERMS, f);
    fwrite(y, sizeof(double), 18 * N, f);
    fwrite(&rng, sizeof rng, 1, f);
    fwrite(&tnow, sizeof tnow, 1, f);
    fwrite(&V, sizeof V, 1, f);
    fwrite(cnt, sizeof cnt, 1, f);
    fclose(f);
    printf("run ok: %d beads, %d natives, t = %.3f, V = %.6f, rebuilds %lld\n", N, nn, tnow, V, (long long)cnt[1]);
    return 0;
}

/* Each malformed input must come back as XMD_ERR_ARG with a message. */
static int errors(void) {
    const double box[3] = {0, 0, 0}, temp = 0.35;
    const uint64_t seed = 1;
    const int nn = build_system();
    xmd_params p;
    xmd_topology t;
    xmd_engine *eng = NULL;
    int bad = 0;
    struct {
        const char *what;
        int kind;
    } cases[] = {{"dt", 0}, {"chain offsets", 1}, {"native order", 2}, {"atype", 3}};
    for (size_t c = 0; c < sizeof cases / sizeof cases[0]; ++c) {
        set_params(&p);
        set_topology(&t, nn, box, &temp, &seed);
        int32_t bad_chain[NCH + 1] = {0, LEN + 5, 2 * LEN - 1};
        int32_t swapped[1] = {0};
        int8_t at_bad[N];
        switch (cases[c].kind) {
        case 0: p.dt = -1.0; break;
        case 1: t.chain_offset = bad_chain; break;
        case 2:
            swapped[0] = nat_i2[0];
            t.nat_i1 = swapped;
            t.nat_i2 = nat_i1;
            t.n_native = 1;
            break;
        case 3:
            memcpy(at_bad, atype, N);
            at_bad[7] = XMD_NUM_AA;
            t.atype = at_bad;
            break;
        }
        xmd_status st = xmd_engine_create(&p, &t, &eng);
        const char *msg = xmd_last_error(NULL);
        if (st != XMD_ERR_ARG || !msg || !*msg) {
            fprintf(stderr, "%s: status %d, message '%s'\n", cases[c].what, (int)st, msg ? msg : "");
            if (st == XMD_OK) xmd_engine_destroy(eng);
            bad = 1;
        } else {
            printf("%s refused: %s\n", cases[c].what, msg);
        }
    }
    if (!bad) printf("errors ok\n");
    return bad;
}

int main(int argc, char **argv) {
    printf("%s\n", xmd_version());
    if (argc == 3 && !strcmp(argv[1], "run")) return run(argv[2]);
    if (argc == 2 && !strcmp(argv[1], "errors")) return errors();
    fprintf(stderr, "usage: %s run <out.bin> | errors\n", argv[0]);
    return 2;
}
