/*
 * cd_oracle.h -- C ABI of the CPU oracle for the NSGA-III codon-design loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU, the
 * generation loop of the reference (src/main.cu:700-732 with the objective
 * evaluators of include/utils.cuh, the operators of include/mutation.cuh and
 * the selection of include/sorting.cuh).  It is the checker that the HIP
 * engine is compared against; nothing in the engine links or calls it.
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * load it.
 *
 * Every function is deterministic.  Where the reference is nondeterministic
 * (cuRAND streams, atomic ordering, spin-lock tie races) the oracle uses the
 * documented deterministic replacement that the engine must reproduce bit for
 * bit (Philox4x32-10 streams, canonical tie rules, ordered niching).
 */
#ifndef CD_ORACLE_H
#define CD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- primitives ------------------------------------------------------- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_det_exp(double x);
int or_dasdennis(int M, int H, float *out /* C(M+H-1,H) x M */);

/* Hairpin stem of one CDS given as nucleotide codes (A=0,C=1,G=2,U=3).
 * mode 0: closed form with the canonical tie rule (returns l,p,q)
 * mode 1: transliteration of the reference's anti-diagonal walk (l only)
 * mode 2: brute force over all innermost pairs (l only)                  */
int or_sl_cds(const uint8_t *codes, int n, int mode, int32_t *l, int32_t *p, int32_t *q);

/* ---- evaluation -------------------------------------------------------- */
/* Canonical evaluation of `count` designs (int8 synonym indices [count][m][L]).
 * obj [count][6], idx [count][12], pql [count][3].                          */
int or_eval(int L, int m, const int8_t *aa, const char *codon_nt, const float *w,
            const float *cps, float gc3_ref, const int8_t *syn, int count,
            float *obj, int8_t *idx, int32_t *pql);

/* Reference-order FP32 CAI / CPB of every CDS (the reference's thread-strided
 * partials plus halving tree, utils.cuh:271-409, with `threads` lanes) and the
 * FP64 closed forms.  out32/out64 are [count][m][2] = (CAI, CPB).           */
int or_cai_cpb_ref(int L, int m, const int8_t *aa, const char *codon_nt, const float *w,
                   const float *cps, const int8_t *syn, int count, int threads,
                   float *out32, double *out64);

/* The mutation of one child (mutationKernel's operator choice and operators,
 * main.cu:172-211 with include/mutation.cuh): child `child` of generation
 * `gen` from its parent's genome [m][L] and record (obj[6], idx[12], pql[3]);
 * out [m][L] receives the mutated genome.                                   */
int or_mutate(int L, int m, const int8_t *aa, const char *codon_nt, const float *w, const float *cps,
              float gc3_ref, float Pm, uint64_t seed, uint32_t gen, uint32_t child, const int8_t *parent,
              const float *pobj, const int8_t *pidx, const int32_t *ppql, int8_t *out);

/* ---- selection ---------------------------------------------------------- */
/* Non-dominated sorting (sorting.cuh:974-1089) of P objective vectors, peeled
 * until at least N are ranked.  front_of[P] = front index or -1.            */
int or_nds(const float *obj, int P, int N, int32_t *front_of, int32_t *n_fronts);

/* One full NSGA-III selection (sortingKernel, main.cu:409-454) on a pool of
 * P = 2N objective vectors.  extremes[36] is read and updated in place.
 * sorted_out[N] receives the survivors; front_of[P]; norm[18] receives
 * ideal[6], worst[6], nadir[6]; flags[2] = (N_cut, hyp_exception).        */
int or_select(const float *obj, int N, int H, uint32_t gen, uint64_t seed,
              float *extremes, int32_t *sorted_out, int32_t *front_of,
              float *norm, int32_t *flags);

/* ---- whole job ---------------------------------------------------------- */
void *or_job_create(int L, int m, int N, int H, float Pm, uint64_t seed, const int8_t *aa,
                    const char *codon_nt, const float *w, const float *cps, float gc3_ref);
int or_job_init(void *job);
int or_job_step(void *job, int gens);
/* genome [2N][m][L] int8, obj [2N][6], idx [2N][12], pql [2N][3],
 * sorted [N], front_of [2N], state: extremes[36], ideal[6], worst[6], nadir[6] */
int or_job_read(void *job, int8_t *genome, float *obj, int8_t *idx, int32_t *pql,
                int32_t *sorted, int32_t *front_of, float *state);
int or_job_generation(void *job);
void or_job_destroy(void *job);
int or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
