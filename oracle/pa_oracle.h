/*
 * pa_oracle.h -- CPU restatement of PhyloAcc-ST's per-element MCMC path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (phyloacc_amd/) never links, loads or calls it.
 *
 * It restates, in plain C and FP64 log space, one site at a time, the
 * reference routines cited per function in pa_oracle.c (src/PhyloAcc-ST/...).
 * Random numbers come from a counter-based Philox4x32-10 stream whose slot
 * layout is fixed (see or_slot_* below) so that the GPU engine, fed the same
 * seed, consumes exactly the same uniforms.
 */
#ifndef PA_ORACLE_H
#define PA_ORACLE_H
#include <stdint.h>

#define OR_MAXN 512

/* leaf codes (shared meaning with the product's encoder) */
enum { OR_A = 0, OR_C = 1, OR_G = 2, OR_T = 3, OR_GAP = 4, OR_OTHER = 5,
       OR_R = 6, OR_Y = 7, OR_K = 8, OR_M = 9, OR_S = 10, OR_W = 11 };

/* chain failure and warning bits (or_chain_out.flags): log-likelihood NaN,
 * +inf or -inf at some state; a Gamma draw stopped at its rejection cap; a
 * rate proposal outside (1e-12, 1e8) skipped (the reference's warning,
 * src/PhyloAcc-ST/bpp_c.cpp:1024-1041) */
enum { OR_FLAG_NAN = 1, OR_FLAG_POS_INF = 2, OR_FLAG_NEG_INF = 4, OR_FLAG_SAMPLER = 8, OR_FLAG_PROPOSAL_RANGE = 16 };

/* RNG stream tags (bits 16..23 of counter word 3) */
enum { OR_TAG_ANC = 1, OR_TAG_Z = 2, OR_TAG_R2 = 3, OR_TAG_R1 = 4, OR_TAG_TRANS = 5, OR_TAG_PI = 6, OR_TAG_WL = 7,
       OR_TAG_GT_BRANCH = 8, OR_TAG_GT_TREE = 9, OR_TAG_GT_HIST = 10 };

typedef struct {
    int burnin, mcmc, adapt_freq;          /* BURNIN, MCMC, ADAPT_FREQ */
    double init_crate, init_nrate;         /* INIT_CONSERVE_RATE, INIT_ACCE_RATE */
    double cprior_a, cprior_b;             /* CONSERVE_PRIOR_A/B (shape, scale) */
    double nprior_a, nprior_b;             /* ACCE_PRIOR_A/B */
    int ropt; double cub, nlb;             /* ROPT, CUB, NLB */
    double init_grate, init_lrate, init_lrate2;
    double hyper_grate_a, hyper_grate_b;
    double hyper_lrate_a, hyper_lrate_b;
    double hyper_lrate2_a, hyper_lrate2_b;
    double gap_prop;                       /* GAP_PROP (missing_thres) */
    double conserve_prop;                  /* CONSERVE_PROP */
    double constomis;                      /* CONSTOMIS */
    double nconstomis;                     /* fixed 1 in the reference driver */
    double trim_gap_percent;               /* TRIM_GAP_PERCENT (revgap) */
    int min_len;                           /* MIN_LEN */
    int prune_tree;                        /* PRUNE_TREE */
    int gapchar;                           /* GAPCHAR (first char) */
    uint64_t seed;                         /* SEED */
    int r1_shared_uniform;                 /* 1: reference r1 quirk (one uniform proposes and accepts) */
} or_options;

typedef struct {
    int S, N;
    int children[OR_MAXN][2];
    int parent[OR_MAXN];
    double dist[OR_MAXN];
    double pi[4];
    double Q[16];                          /* row = from state */
    int upper[OR_MAXN];                    /* BPP::upper membership (outgroup ancestors + root) */
    int upper_conserve[OR_MAXN];           /* BPP::upper_conserve membership */
    int conserve[OR_MAXN]; int n_conserve; /* CONSERVE species ids */
} or_model;

typedef struct or_element or_element;

typedef struct {
    /* traces, length T = burnin + mcmc (caller-owned, may be NULL) */
    double *loglik, *full, *nrate, *crate, *grate, *lrate, *l2rate;
    signed char *Z;                        /* [T*N] */
    double *postZ;                         /* [N*4] Output_init frequencies (may be NULL) */
    double max_loglik; int max_m; int max_Z[OR_MAXN];
    double vb;                             /* Eval2 marginal likelihood */
    double med[5];                         /* Output_init medians: n, c, g, l, l2 */
    int failure;
    int accept_n, accept_c;                /* counters after the chain (carry over) */
    int flags;                             /* OR_FLAG_* bits */
} or_chain_out;

#ifdef __cplusplus
extern "C" {
#endif

/* RNG */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_uniform(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, int lane);
double or_gamma_stream(uint64_t seed, uint32_t c1, uint32_t c2, uint32_t c3, double shape, double scale);
double or_beta_stream(uint64_t seed, uint32_t c1x, uint32_t c1y, uint32_t c2, uint32_t c3, double a, double b);

/* tree / model text format */
int or_parse_mod(const char *text, or_model *m, char *names /* [N*64] */, int max_names);
void or_set_groups(or_model *m, const int *outgroup, int n_out, const int *conserve, int n_cons);

/* transition matrices: logPt[i*4+j] = log P_{j->i}(t*r) */
void or_expm(const double *Q, double t, double *P /* [16], P[i*4+j] = P_{i->j} */);
void or_log_trans(const or_model *m, double t, double *logPt);

/* element encoding (BPP_C constructor) */
or_element *or_encode(const or_model *m, const or_options *o, const char *const *rows, long start, int GG);
void or_free(or_element *e);
int or_elem_filtered(const or_element *e);
int or_elem_info(const or_element *e, int *out /* [8]: GG, root, n_nodes, n_internal, S, N, filtered, 0 */);
void or_elem_arrays(const or_element *e, int *children2 /*[N*2]*/, int *parent2, double *dist2,
                    int *missing, int *nodes, int *internal, int *Tg /*[GG*N]*/, unsigned char *codes /*[GG*S]*/);

/* log-likelihood at fixed Z, rates (bpp_c2.cpp:392-425 form) */
double or_loglik(const or_model *m, or_element *e, const int *Z, double r1, double r2);
/* PhyloAcc-GT base frequencies: the gene-tree log-likelihood at pi = (a,
 * 1/2 - a, 1/2 - a, a) and one sample_pi step (src/PhyloAcc-GT/bpp_c.cpp:
 * 1508-1597); arguments after the first ones as or_gt_loglik. */
double or_gt_loglik_pi(const or_model *m, double a, int S, const int *gchildren, int K, const int *seg_gene,
                       const int *seg_sp, const double *seg_len, const int *Z, double r1, double r2,
                       const char *const *rows, long start, int GG, int gapchar, double gap_prop);
int or_gt_sample_pi(const or_model *m, uint64_t seed, int sweep, double pi_delta, double prior_a, double prior_b,
                    int resample, double *piA, double *loglik, int S, const int *gchildren, int K,
                    const int *seg_gene, const int *seg_sp, const double *seg_len, const int *Z, double r1,
                    double r2, const char *const *rows, long start, int GG, int gapchar, double gap_prop);
/* One Wang-Landau indicator step of PhyloAcc-GT's Gibbs; st[7] as pa_wl_step. */
int or_wl_step(uint64_t seed, int m, const double *loglik, const double *logpz, double *st);
/* gBGC model (PhyloAcc-ST-GBGC): Q(s, b) from Q0, and the log-likelihood with
 * per-branch classes (Z, B) and sel = {conserved s, accelerated s, bias}. */
void or_gbgc_rate(const double *Q0, double s, double b, double *Q);
double or_loglik_gbgc(const or_model *m, or_element *e, const int *Z, const int *B, const double *sel);

/* full three-model chain of one element, outs[0..2] = M0, M1, M2 (run order resZ 0, 2, 1) */
int or_run_element(const or_model *m, const or_options *o, or_element *e, or_chain_out *outs);

/* gene-tree log-likelihood at fixed species-branch Z and rates (PhyloAcc-GT pruning) */
double or_gt_loglik(const or_model *m, int S, const int *gchildren, int K, const int *seg_gene, const int *seg_sp,
                    const double *seg_len, const int *Z, double r1, double r2, const char *const *rows, long start,
                    int GG, int gapchar, double gap_prop);

/* PhyloAcc-GT's species tree (BPP::InitPhyloTree, src/PhyloAcc-GT/bpp.cpp:40-84):
 * parent [N] (N at the root), node heights [N], population sizes theta [N]. */
typedef struct { int S, N; const int *parent; const double *height; const double *theta; } or_sptree;
/* A gene tree (GTree) as topology + heights: gch [N*2] (-1 at leaves; leaf i
 * is species i), gpar [N] (N at the root), gh [N], gmiss [N] (missing_gene).
 * Everything GTree caches per species (parent_gene2, temp_coal) follows from
 * these and the species tree. */
int or_gt_species(const or_sptree *sp, const int *gch, const double *gh, int v);
double or_gt_prior(const or_sptree *sp, const int *gch, const int *gpar, const double *gh, int groot);
double or_gt_logpz(const or_sptree *sp, const int *gch, const double *gh, const int *gmiss, const int *Z,
                   double c2m, double n2m);
int or_gt_pieces(const or_sptree *sp, const int *gch, const int *gpar, const double *gh, int groot, int *seg_gene,
                 int *seg_sp, double *seg_len);
void or_gt_site_logliks(const or_model *m, int S, const int *gchildren, int groot, int K, const int *seg_gene,
                        const int *seg_sp, const double *seg_len, const int *Z, double r1, double r2,
                        const char *const *rows, long start, int GG, int gapchar, double gap_prop, double *site);
/* One GTree::Sample_BranchLen move; returns 1 accepted, 0 not, -1 error. */
int or_gt_branch_len(const or_model *m, double piA, const or_sptree *sp, const int *gch, const int *gpar, double *gh,
                     const int *gmiss, int groot, const int *Z, double c_rate, double n_rate, double c2m, double n2m,
                     int gnode, double delta, int indicator, const int *lens, double *loglik, double *logpz,
                     double *prior, uint64_t seed, int m_sweep, int slot, const char *const *rows, long start, int GG,
                     int gapchar, double gap_prop);

/* GTree::Update_Tg(tosample = true): the history Tg [N * len] of sites g < len. */
void or_gt_sample_history(const or_model *m, double piA, const or_sptree *sp, const int *gch, const int *gpar,
                          const double *gh, int groot, const int *Z, double c_rate, double n_rate, uint64_t seed,
                          int m_sweep, int len, const char *const *rows, long start, int GG, int gapchar,
                          double gap_prop, signed char *Tg);
/* One GTree::Sample_tree2 move; returns 1 accepted, 0 rejected. */
int or_gt_sample_tree(const or_model *m, double piA, const or_sptree *sp, int *gch, int *gpar, double *gh,
                      int *groot_io, const int *Z, double c_rate, double n_rate, double c2m, double n2m, int root_ingrp,
                      int branch, int indicator, const int *lens, const signed char *Tg, double *loglik,
                      double *logpz, uint64_t seed, int m_sweep, int slot, const char *const *rows, long start, int GG,
                      int gapchar, double gap_prop);

/* test hook: or_gt_sample_tree with the complete Hastings ratio (exact) and the candidates' scores */
int or_test_sample_tree(const or_model *m, double piA, const or_sptree *sp, int *gch, int *gpar, double *gh,
                        int *groot_io, const int *Z, double c_rate, double n_rate, double c2m, double n2m,
                        int root_ingrp, int branch, int indicator, const int *lens, const signed char *Tg,
                        double *loglik, double *logpz, uint64_t seed, int m_sweep, int slot, const char *const *rows,
                        long start, int GG, int gapchar, double gap_prop, int exact, int *cand_node,
                        double *cand_score, int *n_cand);

/* test hook: one Z forward-backward draw on a bare tree */
int or_test_update_Z(int N, const int *nodes, int n_nodes, const int *parent, const int *fixZ, const double *E,
                     double g, double l, double l2, uint64_t seed, int m, int ms, int *Zout);

/* test hook: Eval2 over given traces of model slot ms (0 = M0, 1 = M1, 2 = M2) */
double or_test_eval2(const or_model *m, const or_options *o, or_element *e, int ms, const double *full,
                     const double *nrate, const double *crate, const signed char *Z, const int *max_Z);

/* test hook: one site's ancestral draw at fixed Z and rates, uniforms of sweep m_sweep */
int or_test_sample_site(const or_model *m, or_element *e, const int *Z, double r1, double r2, uint64_t seed,
                        int m_sweep, int ms, int site, int *states);

/* CPU baseline: many elements, OpenMP guided schedule; returns sweeps done */
long or_run_batch(const or_model *m, const or_options *o, or_element **es, int n, int nthreads,
                  double *max_loglik /* [n*3] */);

#ifdef __cplusplus
}
#endif
#endif
