/*
 * phyloacc_gpu.h -- C ABI of the MI355X engine for PhyloAcc-ST's per-element
 * MCMC (Felsenstein pruning + Gibbs Z updates + Metropolis r1/r2 updates).
 *
 * Plain C types only (no C++ and no torch types cross this boundary).  Every
 * call returns a pa_status; PA_OK is 0.  pa_last_error() gives the message
 * of the last failing call on an engine.  No exceptions cross the ABI.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference repository).
 */
#ifndef PHYLOACC_GPU_H
#define PHYLOACC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PA_OK = 0,
    PA_E_INPUT = 1,     /* malformed input (tree, options, alignment, ids) */
    PA_E_NUMERIC = 2,   /* numeric failure (NaN likelihood, bad eigensystem) */
    PA_E_DEVICE = 3,    /* HIP runtime error or no usable gfx950 device */
    PA_E_OOM = 4,       /* device memory exhausted / over the HBM budget */
    PA_E_STATE = 5      /* call not valid in the engine's current state */
} pa_status;

enum { PA_FP64 = 0, PA_FP32 = 1 };                 /* partial-likelihood precision */
enum { PA_RUN_BATCHED = 0, PA_RUN_PER_ELEMENT = 1 }; /* launch strategy */
enum { PA_ELEM_OK = 0, PA_ELEM_FILTERED = 1, PA_ELEM_NUMERIC_FAIL = 2 };

/* Per-chain failure and warning bits (pa_chain_result.flags).  The reference
 * marks an element failed when its marginal likelihood is NaN
 * (src/PhyloAcc-ST/bpp_c2.cpp:339-348) and warns on out-of-range rate
 * proposals (bpp_c.cpp:1024-1041); here every chain records: */
enum {
    PA_FLAG_NAN = 1,            /* the log-likelihood was NaN at some state (status NUMERIC_FAIL) */
    PA_FLAG_POS_INF = 2,        /* ... +inf at some state (status NUMERIC_FAIL) */
    PA_FLAG_NEG_INF = 4,        /* ... -inf at some state: data impossible at that state's rates;
                                   status NUMERIC_FAIL when the chain's maximum is not finite */
    PA_FLAG_SAMPLER = 8,        /* a Gamma draw stopped at its rejection cap (status NUMERIC_FAIL) */
    PA_FLAG_PROPOSAL_RANGE = 16 /* a rate proposal fell outside (1e-12, 1e8) and was skipped (warning) */
};

/* leaf codes produced by the encoder */
enum { PA_CODE_A = 0, PA_CODE_C = 1, PA_CODE_G = 2, PA_CODE_T = 3, PA_CODE_GAP = 4,
       PA_CODE_OTHER = 5, PA_CODE_R = 6, PA_CODE_Y = 7, PA_CODE_K = 8, PA_CODE_M = 9,
       PA_CODE_S = 10, PA_CODE_W = 11 };

/* Run parameters, one field per reference parameter-file key
 * (src/PhyloAcc-ST/main.cpp:134-277, defaults main.cpp:35-90). */
typedef struct {
    int32_t burnin;            /* BURNIN (200) */
    int32_t mcmc;              /* MCMC (800) */
    int32_t adapt_freq;        /* ADAPT_FREQ (500) */
    double init_crate;         /* INIT_CONSERVE_RATE (0.5) */
    double init_nrate;         /* INIT_ACCE_RATE (1) */
    double cprior_a, cprior_b; /* CONSERVE_PRIOR_A/B (5, 0.04) */
    double nprior_a, nprior_b; /* ACCE_PRIOR_A/B (10, 0.2) */
    int32_t ropt;              /* ROPT (1) */
    double cub, nlb;           /* CUB (1), NLB (0.6) */
    double init_grate, init_lrate, init_lrate2;   /* INIT_GRATE/LRATE/LRATE2 (0.5, 0.3, 0.1) */
    double hyper_grate_a, hyper_grate_b;          /* HYPER_GRATE_A/B (3, 1) */
    double hyper_lrate_a, hyper_lrate_b;          /* HYPER_LRATE_A/B (1, 9) */
    double hyper_lrate2_a, hyper_lrate2_b;        /* HYPER_LRATE2_A/B (1, 9); A = 0 -> Dollo */
    double gap_prop;           /* GAP_PROP (0.8) */
    double conserve_prop;      /* CONSERVE_PROP (0.8) */
    double constomis;          /* CONSTOMIS (0.01) */
    double trim_gap_percent;   /* TRIM_GAP_PERCENT (1 = off) */
    int32_t min_len;           /* MIN_LEN (50) */
    int32_t prune_tree;        /* PRUNE_TREE (0) */
    int32_t gapchar;           /* GAPCHAR ('-') */
    uint64_t seed;             /* SEED (5) */
    int32_t r1_shared_uniform; /* 1 = the reference's r1 step (one uniform proposes and accepts,
                                  src/PhyloAcc-ST/bpp_c.cpp:993,1010,1078); 0 = fresh uniform */
    int32_t compress_patterns; /* 1 = prune over unique site patterns (default); 0 = one per site */
} pa_options;

/* Fills o with the reference defaults. */
void pa_default_options(pa_options *o);

/* Tree + substitution model in the reference's node numbering
 * (src/PhyloAcc-ST/newick.cpp:382-428; bpp.cpp:84-117): leaves 0..S-1,
 * internal S..N-1 in post-order, root N-1, children[i][0] < children[i][1]. */
typedef struct {
    int32_t S;                 /* species (leaves); N = 2S-1 */
    const int32_t *children;   /* [N*2], -1 for leaves */
    const int32_t *parent;     /* [N], -1 for the root */
    const double *dist;        /* [N] branch length above each node */
    const double *pi;          /* [4] BACKGROUND */
    const double *Q;           /* [16] RATE_MAT, row = from state */
    const int32_t *outgroup;   int32_t n_outgroup;   /* OUTGROUP species ids */
    const int32_t *conserve;   int32_t n_conserve;   /* CONSERVE species ids */
} pa_model;

typedef struct pa_engine pa_engine;

/* Parsed .mod file.  Replaces LoadPhyloTree (src/PhyloAcc-ST/newick.cpp:431-517):
 * same numbering, float-rounded branch lengths, names of every node.
 * names receives N NUL-terminated strings of at most 63 chars at stride 64. */
pa_status pa_parse_mod(const char *text, int32_t max_nodes, int32_t *S, int32_t *children,
                       int32_t *parent, double *dist, double *pi, double *Q, char *names);

/* Engine over one tree + model + options on device `device`.
 * Replaces the BPP constructor (src/PhyloAcc-ST/bpp.hpp:173-404: eigensystem of Q,
 * outgroup/conserve ancestor sets) and BPP::InitMCMC (bpp.cpp:186-224). */
pa_status pa_engine_create(const pa_model *model, const pa_options *opts, int32_t device,
                           pa_engine **out);
void pa_engine_destroy(pa_engine *e);
const char *pa_last_error(const pa_engine *e);

/* Encodes elements from an alignment whose S rows are in tree leaf order
 * (BPP::MatchProfAndTree, bpp.cpp:133-184, is the caller's job).  Element i
 * is columns [starts[i], starts[i]+lengths[i]) of every row.  Replaces the
 * BPP_C constructor (src/PhyloAcc-ST/bpp_c.hpp:143-476): encoding, trimming,
 * missing species, CONSERVE filter, outgroup-lineage collapse.  Writes one
 * engine element id and one PA_ELEM_* status per element. */
pa_status pa_elements_add(pa_engine *e, const char *const *rows, const int64_t *starts,
                          const int32_t *lengths, int32_t n, int32_t *ids, int32_t *status);

/* Elements that carry their own tree over the same S leaves: the per-element
 * gene trees of PhyloAcc-GT (src/PhyloAcc-GT/genetree.hpp:16-127).  children:
 * [n*N*2], internal ids in post-order, root N-1.  Encoding as pa_elements_add
 * without the outgroup collapse and the CONSERVE filter. */
pa_status pa_elements_add_trees(pa_engine *e, const char *const *rows, const int64_t *starts,
                                const int32_t *lengths, int32_t n, const int32_t *children, int32_t *ids,
                                int32_t *status);

/* Substitution length of each gene branch from its pieces in species branches
 * (K pieces: gene node, species branch, length) at rates 1 / r1 / r2 for
 * species-branch Z = 0 / 1 / 2.  Replaces the per-piece getlogTMc composition
 * of GTree::Update_Tg (src/PhyloAcc-GT/genetree.cpp:1330-1426;
 * src/PhyloAcc-GT/bpp.cpp:221-235): P(a)P(b) = P(a+b) for one Q, and pieces
 * shorter than 1e-8 are the identity there and add nothing here. */
pa_status pa_gene_lengths(int32_t N, int32_t K, const int32_t *seg_gene, const int32_t *seg_species,
                          const double *seg_len, const int32_t *Z, double r1, double r2, double *out);

/* Host-side view of one encoded element (for checks and drop-in outputs). */
typedef struct {
    int32_t filtered, GG, P, root, n_nodes, n_internal;
} pa_element_info;
pa_status pa_element_get(const pa_engine *e, int32_t id, pa_element_info *info,
                         int32_t *children2 /*[N*2] or NULL*/, int32_t *parent2 /*[N]*/,
                         double *dist2 /*[N]*/, int32_t *missing /*[N]*/, int32_t *nodes /*[N]*/,
                         uint8_t *site_codes /*[GG*S] or NULL*/, int32_t *site_pattern /*[GG]*/,
                         int32_t *weights /*[P]*/);

/* Per-element log-likelihood at fixed Z and rates, on the device.
 * Replaces BPP_C::log_f_Xz(log_pi, num_base, Z, caches) (src/PhyloAcc-ST/bpp_c2.cpp:392-425)
 * evaluated at log_cache_TM_cons/neut built for r1/r2 (bpp_c.cpp:187-228).
 * Z: [n*N] int32 (0 neutral, 1 conserved, 2 accelerated, -1 outside the element's tree);
 * rates: [n*2] (r1 conserved, r2 accelerated).  out: [n]. */
pa_status pa_loglik(pa_engine *e, const int32_t *ids, int32_t n, const int32_t *Z,
                    const double *rates, int32_t precision, double *out);

/* Per-element log-likelihood of the element's own tree with per-call branch
 * lengths [n*N] (all at rate 1): batched pruning over uneven gene trees. */
pa_status pa_loglik_lengths(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths,
                            int32_t precision, double *out);

/* Gene-tree log-likelihoods as pa_loglik_lengths, each element with its own
 * strand-symmetric base frequencies pi = (a, 1/2 - a, 1/2 - a, a), piA[n] = a
 * in (0, 1/2): the element's Q is BPP_C::getQmat of pi and the model's
 * exchangeabilities (src/PhyloAcc-GT/bpp_c.cpp:160-185; exchangeabilities as
 * in src/PhyloAcc-GT/bpp.hpp:346-353), rebuilt and eigendecomposed on the
 * device, and the root sum uses pi. */
pa_status pa_loglik_lengths_pi(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths,
                               const double *piA, int32_t precision, double *out);

/* One BPP_C::sample_pi step (src/PhyloAcc-GT/bpp_c.cpp:1508-1597) for n
 * gene-tree elements at their lengths [n*N], sweep m, over the full element
 * (no Wang-Landau split).  piA[n] and loglik[n] hold each element's current
 * a and log-likelihood and are updated in place; accepted[n] is 1 when the
 * element's step was accepted (an unchanged proposal counts as accepted, as
 * in the reference).  resample = 0: a uniform step of half-width pi_delta
 * clamped to [0.05, 0.45] with the reference's boundary terms and a
 * Metropolis decision under the Beta(prior_a, prior_b) prior on 2a;
 * resample = 1: a = Beta(prior_a + #A + #T, prior_b + #C + #G) / 2 over the
 * element's leaf bases, always taken (bpp_c.cpp:1520-1521,1557-1566).  The
 * proposals' likelihoods are one batched launch. */
pa_status pa_sample_pi(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths, int32_t m,
                       double pi_delta, double prior_a, double prior_b, int32_t resample, int32_t precision,
                       double *piA, double *loglik, int32_t *accepted);

/* Gene-tree log-likelihoods over two prefixes of each element's sites: out0
 * sums its first lens0[i] sites, out1 its first lens1[i] (0 <= lens1 <= lens0
 * <= the element's sites), the two log_lik entries of the Wang-Landau blocks
 * of PhyloAcc-GT's Gibbs (src/PhyloAcc-GT/bpp_c.cpp:359-380,436-443); one
 * launch prunes each element once.  piA: NULL for the model's base
 * frequencies, or per element as in pa_loglik_lengths_pi. */
pa_status pa_loglik_lengths_split(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths,
                                  const double *piA, const int32_t *lens0, const int32_t *lens1,
                                  int32_t precision, double *out0, double *out1);

/* One Wang-Landau indicator step (src/PhyloAcc-GT/bpp_c.cpp:399-440) for n
 * elements at sweep m, on the host.  loglik, logpz: [n*2] (the two prefixes'
 * log-likelihoods and missing-pattern terms); state: [n*7] = two log
 * normalising constants, two momenta, the learning rate, two visit counts,
 * set from loglik + logpz at m = 0 and updated in place; indicator: [n], the
 * distribution (0 or 1) the next sweep targets. */
pa_status pa_wl_step(pa_engine *e, int32_t n, int32_t m, const double *loglik, const double *logpz,
                     double *state, int32_t *indicator);

/* ---- PhyloAcc-GT gene trees inside the species tree ----
 * The species tree of PhyloAcc-GT (BPP::InitPhyloTree, src/PhyloAcc-GT/
 * bpp.cpp:40-84, with the population sizes of src/PhyloAcc-GT/main.cpp:
 * 379-413): parent [N] (N at the root), node heights [N] (only differences
 * matter) and theta [N] (internal species; the coalescence rate of k lineages
 * in species s is k(k-1)/theta_s).  A gene tree (GTree, src/PhyloAcc-GT/
 * genetree.hpp:16-127) is passed as children [N*2] (-1 at leaves; leaf i is
 * species i), parent [N] (N at the gene root) and height [N].  Whatever GTree
 * caches per species (parent_gene2, temp_coal) follows from these: a gene
 * node lies in the lowest species on the climb from its first leaf whose
 * parent species is not lower than the node (GTree::findSp, genetree.cpp:
 * 2521-2534), and a gene lineage crosses every species from its node's up to
 * its parent's. */
typedef struct {
    int32_t S, N;
    const int32_t *parent;
    const double *height;
    const double *theta;
} pa_species_tree;

/* GTree::priorTree (src/PhyloAcc-GT/genetree.cpp:560-584): the multispecies-
 * coalescent log density of the gene tree's coalescence times.  Host only. */
pa_status pa_gt_prior(const pa_species_tree *sp, const int32_t *children, const int32_t *parent,
                      const double *height, double *out);

/* GTree::get_logpZ1 (src/PhyloAcc-GT/genetree.cpp:1156-1190): sum of
 * log P(missing | Z) over the leaves and over the gene nodes born in internal
 * species, P(missing | Z = 1) = cons_to_mis, P(missing | Z != 1) =
 * ncons_to_mis; a gene node is missing when both children are.  leaf_missing
 * [S], Z [N].  Host only. */
pa_status pa_gt_logpz(const pa_species_tree *sp, const int32_t *children, const int32_t *parent,
                      const double *height, const int32_t *leaf_missing, const int32_t *Z, double cons_to_mis,
                      double ncons_to_mis, double *out);

/* Substitution length of every gene branch (lengths [N], 0 at the root): its
 * pieces in the species it crosses at each species' rate (Z 2: r2
 * accelerated, Z 1: r1 conserved, else 1); pieces shorter than 1e-8 add
 * nothing.  The one-matrix-per-piece composition of GTree::Update_Tg
 * (src/PhyloAcc-GT/genetree.cpp:1330-1426) is P of the summed length.  These
 * are the lengths pa_loglik_lengths takes.  Host only. */
pa_status pa_gt_lengths(const pa_species_tree *sp, const int32_t *children, const int32_t *parent,
                        const double *height, const int32_t *Z, double r1, double r2, double *lengths);

/* One GTree::Sample_BranchLen move (src/PhyloAcc-GT/genetree.cpp:1658-1965)
 * for each of n gene-tree elements, the proposals' likelihoods in one launch.
 * Element i moves gene node gnode[i] of its gene tree (gchildren [n*N*2],
 * gparent [n*N], gheight [n*N]; the tree must be the element's stored tree,
 * pa_elements_add_trees): the height is drawn uniformly in the window
 * h +- delta * theta / 2 of the node's species, clipped below at the bottom
 * of the lowest species both children's lineages cross or at the taller
 * child, above at the parent gene node when that is born in the same
 * species, else at the top of the species (the gene root: 10 theta_root / 2,
 * or h + 5 theta_root / 2 when higher); then Metropolis-Hastings on the
 * prefix log-likelihood of indicator[i] (sites g < lens[2i + indicator[i]];
 * loglik[2i + 1] sums g < lens[2i + 1] as well), the missing-gene term when
 * logpz[2i + indicator[i]] != 0, the coalescent prior and the clipped window's
 * proposal ratio.  Z [n*N] species classes, rates [n*2] (r1 conserved, r2
 * accelerated), piA [n] or NULL as pa_loglik_lengths_pi.  On acceptance
 * gheight, loglik [n*2], logpz [n*2] (both entries, or entry 0 while
 * logpz[2i + 1] == 0) and prior [n] take the proposal's values and
 * accepted[i] = 1.  An empty window draws nothing and is not accepted.
 * Draws: tag GT_BRANCH, sub-stream slot, sweep m; the height's uniform
 * (floored at 1e-5) at 0, the decision's at 1. */
pa_status pa_gt_branch_len(pa_engine *e, const pa_species_tree *sp, const int32_t *ids, int32_t n,
                           const int32_t *gchildren, const int32_t *gparent, double *gheight, const int32_t *Z,
                           const double *rates, const double *piA, const int32_t *gnode, const int32_t *indicator,
                           const int32_t *lens, double delta, double cons_to_mis, double ncons_to_mis, int32_t m,
                           int32_t slot, int32_t precision, double *loglik, double *logpz, double *prior,
                           int32_t *accepted);

/* GTree::Update_Tg with tosample = true (src/PhyloAcc-GT/genetree.cpp:
 * 1416-1532) for n gene-tree elements in one launch: after the pruning pass
 * at the gene trees' heights (gchildren, gparent, gheight as
 * pa_gt_branch_len; Z, rates, piA too), sites g < len[i] get the base of
 * every gene node.  The histories stay in the engine's device memory for the
 * next pa_gt_sample_tree of the same ids (Tg NULL there); Tg, when not NULL,
 * also receives them at Tg[(i * N + v) * stride + g]: leaves their data (0-3
 * a/c/g/t, -1 an ambiguity code, 4 no base), the gene root drawn from
 * pi * L_root, every other internal gene node from P(parent's base -> .) * L
 * over its whole branch (the law of the reference's steps through species
 * boundaries); 4 for a missing gene or a subtree with no base at the site.
 * Draws: tag GT_HIST, sub-stream g, sweep m, slot = node id. */
pa_status pa_gt_sample_history(pa_engine *e, const pa_species_tree *sp, const int32_t *ids, int32_t n,
                               const int32_t *gchildren, const int32_t *gparent, const double *gheight,
                               const int32_t *Z, const double *rates, const double *piA, const int32_t *len,
                               int32_t m, int32_t precision, int8_t *Tg, int32_t stride);

/* One GTree::Sample_tree2 move (src/PhyloAcc-GT/genetree.cpp:2138-2518, with
 * Remove_branch :737-808 and Graft :810-923) for each of n gene-tree elements
 * (distinct ids), the unconstrained candidate set (no deep-coalescence
 * species, the default branch-sampling threshold).  Gene lineage branch[i]
 * is pruned with its parent bp; walking up the species from branch's, every
 * gene node born above branch is a candidate new grandparent at substitution
 * distance (rates of the species crossed), with a candidate above the gene
 * root at 10 theta_root / 2 (or h(root) + 5 theta_root / 2) in the species
 * root; the walk stops after species root_ingrp once a candidate exists.
 * Each candidate is scored over the sites g < lens[2i + indicator[i]] with
 * the history Tg (as pa_gt_sample_history writes it, same stride; NULL: the
 * device histories of the last pa_gt_sample_history, which must have had
 * the same ids in the same order, else PA_E_STATE) and
 * branch's partial, all in one launch; a candidate is drawn with probability
 * proportional to exp(score), then the new sibling (one of its two
 * children: u < 1/2 the first) and bp's height uniformly from where both
 * lineages first share a species up to the candidate.  A second launch
 * scores the regrafted trees; Metropolis-Hastings on the prefix likelihood,
 * the missing-gene term (when logpz[2i + indicator[i]] != 0), the
 * coalescent prior, the two uniform heights' densities and the candidate
 * probabilities of the new and the old grandparent (the reference's ratio,
 * which leaves out the 1/2 of the sibling choice).  An old grandparent that
 * is no candidate rejects the move.  On acceptance gchildren, gparent and
 * gheight take the new tree, which also becomes the element's stored tree,
 * and loglik / logpz update as in pa_gt_branch_len.  Draws: tag GT_TREE,
 * sub-stream slot, sweep m: candidate 0, sibling 1, height 2 (floored at
 * 1e-5), decision 3. */
pa_status pa_gt_sample_tree(pa_engine *e, const pa_species_tree *sp, const int32_t *ids, int32_t n,
                            int32_t *gchildren, int32_t *gparent, double *gheight, const int32_t *Z,
                            const double *rates, const double *piA, const int32_t *branch, const int32_t *indicator,
                            const int32_t *lens, const int8_t *Tg, int32_t stride, double cons_to_mis,
                            double ncons_to_mis, int32_t root_ingrp, int32_t m, int32_t slot, int32_t precision,
                            double *loglik, double *logpz, int32_t *accepted);

/* A device-resident batch of gene trees for the node-height moves: the trees,
 * their classes and rates, base frequencies (piA may be NULL) and chain state
 * (loglik [n*2], logpz [n*2], prior [n], as pa_gt_branch_len keeps them) are
 * copied to HBM once; each pa_gt_session_branch_len is then one
 * GTree::Sample_BranchLen move for every element with the proposal, the
 * likelihood launch and the decision all on the device (one wave per element
 * for the move's logic, the same gene-tree code as pa_gt_branch_len), so a
 * sweep of moves needs no host round trip; pa_gt_session_read returns the
 * heights and state.  The trees must be the elements' stored trees and keep
 * their topology while the session lives (a regraft needs a new session).
 * Moves, draws, decisions, heights, prior and missing-gene term are
 * pa_gt_branch_len's, bit for bit; accepted (n, or NULL) receives the
 * decisions.  Species trees up to N = 507 nodes (an LDS copy per element);
 * larger ones give PA_E_INPUT.  The session keeps every element's partials
 * on the device (two slots per internal node: 2 (N - S) Ppad 20 bytes per
 * element in FP32, 40 in FP64), so a move prunes only the path from the
 * moved node to the root; not enough device memory gives PA_E_OOM. */
typedef struct pa_gt_session pa_gt_session;
pa_status pa_gt_session_create(pa_engine *e, const pa_species_tree *sp, const int32_t *ids, int32_t n,
                               const int32_t *gchildren, const int32_t *gparent, const double *gheight,
                               const int32_t *Z, const double *rates, const double *piA, const double *loglik,
                               const double *logpz, const double *prior, double cons_to_mis, double ncons_to_mis,
                               pa_gt_session **out);
pa_status pa_gt_session_branch_len(pa_gt_session *s, const int32_t *gnode, const int32_t *indicator,
                                   const int32_t *lens, double delta, int32_t m, int32_t slot, int32_t precision,
                                   int32_t *accepted);
/* k height moves in a row, e.g. GT's Gibbs step at every internal gene node
 * (GT bpp_c.cpp:640-660): move j moves gene node gnodes[j*n + i] of element
 * i with draw slot slot0 + j, exactly as k pa_gt_session_branch_len calls
 * with slots slot0, slot0 + 1, ... would; all k moves are queued on the
 * device at once and the call waits once.  accepted: [k*n] or NULL;
 * kernel_ms (or NULL): the summed time of the k likelihood launches, whose
 * counted bytes pa_loglik_last_stats then reports in total. */
pa_status pa_gt_session_sweep(pa_gt_session *s, const int32_t *gnodes, int32_t k, const int32_t *indicator,
                              const int32_t *lens, double delta, int32_t m, int32_t slot0, int32_t precision,
                              int32_t *accepted, double *kernel_ms);
pa_status pa_gt_session_read(pa_gt_session *s, double *gheight, double *loglik, double *logpz, double *prior);
void pa_gt_session_destroy(pa_gt_session *s);

/* Per-element log-likelihood under the gBGC substitution model, on the device.
 * Replaces the pruning of PhyloAcc-ST-GBGC over log_cache_TM_{null,cons,neut}
 * and their _B twins (src/PhyloAcc-ST-GBGC/SRC/bpp_c.cpp:1563-1621), whose
 * rate matrices are BPP_C::getRate(submat, selection, bias) (bpp_c.cpp:13-29)
 * for the six classes of bpp_c.cpp:191-197 and whose transitions are
 * getTransition (bpp_c.cpp:1405-1427) over the element's branch lengths.
 * Branch s takes class Z (0 background: selection 0; 1 conserved: sel[0];
 * 2 accelerated: sel[1]) with bias sel[2] when B[s] != 0, else bias 0.
 * Z, B: [n*N] int32; sel: [n*3] (conserved selection, accelerated selection,
 * gBGC bias).  A class whose Q(s, b) has no time-reversible eigensystem
 * gives NaN.  out: [n]. */
pa_status pa_loglik_gbgc(pa_engine *e, const int32_t *ids, int32_t n, const int32_t *Z, const int32_t *B,
                         const double *sel, int32_t precision, double *out);

/* pa_loglik and pa_loglik_lengths keep their device scratch (workgroup slots
 * and per-call length buffer) for the next call when it is at most 8 GiB;
 * this frees it now, with the pinned host buffers that stage the lengths and
 * the regrafted trees' node tables (no reference counterpart: the reference
 * allocates per element on the host).  Call it before a large pa_run_chains or
 * pa_serve_create when device memory is tight. */
pa_status pa_release_scratch(pa_engine *e);

/* pa_loglik without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  ids [n], Z [n*N], rates
 * [n*2] and out [n] are DEVICE pointers, in the caller's order; the kernel
 * scores an id outside the engine's elements, or a filtered one, as NaN.
 * Scratch is the library's static device array of 256 MiB per GPU, every
 * workgroup slot sized for the engine's longest element: an engine whose
 * longest element needs more than the array holds is rejected with
 * PA_E_INPUT.  The arenas must already be on the device (a pa_loglik or
 * pa_run_chains call after the last pa_elements_add), else PA_E_STATE.  The
 * launch is asynchronous on `stream` (a hipStream_t; NULL: the engine's
 * stream): synchronise it before reading out.  All calls on one GPU share the
 * array, so calls that overlap in time on one GPU race. */
pa_status pa_loglik_noalloc(pa_engine *e, const int32_t *ids, int32_t n, const int32_t *Z, const double *rates,
                            int32_t precision, double *out, void *stream);

/* Results of one model chain of one element (BPP_C::Gibbs bookkeeping,
 * src/PhyloAcc-ST/bpp_c.cpp:470-485). */
typedef struct {
    double max_loglik;          /* log_liks_Z[resZ] */
    int32_t max_m;              /* Max_m */
    int32_t status;             /* PA_ELEM_* */
    double final_nrate, final_crate, final_grate, final_lrate, final_lrate2;
    int32_t accept_n, accept_c; /* adaptive-MH counters at chain end */
    int32_t flags;              /* PA_FLAG_* bits */
    int32_t pad;
} pa_chain_result;

/* Optional device-to-host trace copy, per element and model, T = BURNIN+MCMC. */
typedef struct {
    double *loglik, *full, *nrate, *crate, *grate, *lrate, *l2rate; /* [n*3*T] or NULL */
    int8_t *Z;                                                       /* [n*3*T*N] or NULL */
    int8_t *max_Z;                                                   /* [n*3*N] or NULL */
} pa_traces;

typedef struct {
    int32_t precision;          /* PA_FP64 / PA_FP32 */
    int32_t mode;               /* PA_RUN_BATCHED / PA_RUN_PER_ELEMENT */
    int32_t block_threads;      /* 0 = autotuned */
    int32_t keep_traces;        /* keep full traces on the device (required for pa_traces) */
} pa_run_config;

/* Launch statistics of one pa_run_chains call. */
typedef struct {
    double elapsed_ms;          /* device time of the chain launches (HIP events on the engine stream) */
    double algorithmic_bytes;   /* partial/code/weight bytes the passes must move (counted on device) */
    int32_t grid, block_threads;
    int32_t lds_bytes;          /* LDS per workgroup */
    int32_t n_launches;
    double scratch_bytes;       /* device scratch of the launch: workgroup slots + long elements' own regions */
} pa_run_stats;

/* Runs the three model chains M0, M1, M2 (reference order resZ 0, 2, 1,
 * src/PhyloAcc-ST/main.cpp:466-494) of every listed element on the device.
 * Replaces the OpenMP element loop body (main.cpp:452-500: BPP_C::initMCMC +
 * BPP_C::Gibbs per model).  results: [n*3] in run order (M0, M1, M2).
 * stats may be NULL. */
pa_status pa_run_chains(pa_engine *e, const int32_t *ids, int32_t n, const pa_run_config *cfg,
                        pa_chain_result *results, pa_traces *traces, pa_run_stats *stats);

/* Launch statistics of the last pa_loglik / pa_loglik_lengths / pa_loglik_gbgc
 * call (no reference counterpart): the likelihood kernel's HIP-event duration
 * and its counted algorithmic HBM bytes (partials, exponents, leaf codes,
 * weights), grid, block size, LDS and scratch bytes. */
pa_status pa_loglik_last_stats(pa_engine *e, pa_run_stats *out);

/* A prepared batched chain run, captured once as a HIP graph (no reference
 * counterpart: the reference loops over elements, src/PhyloAcc-ST/main.cpp:452-500).
 * pa_batch_create schedules the ids, plans and allocates the device scratch
 * and result buffers, and captures the launch sequence -- a kernel clearing
 * the counters, the persistent chain kernel, the copies of the result rows
 * into pinned host memory -- on the engine stream.  pa_batch_run replays the
 * graph and writes results ([n*3], like pa_run_chains) and stats; the chains
 * draw from the same streams, so every replay equals pa_run_chains on the
 * same ids.  cfg->mode must be PA_RUN_BATCHED; with cfg->keep_traces the
 * traces are written on the device (as pa_run_chains does) and not returned.  After pa_elements_add the
 * arenas move and pa_batch_run returns PA_E_STATE (prepare the batch again).
 * pa_engine_destroy frees the engine's remaining batches; do not call
 * pa_batch_destroy on them afterwards. */
typedef struct pa_batch pa_batch;
pa_status pa_batch_create(pa_engine *e, const int32_t *ids, int32_t n, const pa_run_config *cfg, pa_batch **out);
pa_status pa_batch_run(pa_batch *b, pa_chain_result *results, pa_run_stats *stats);
void pa_batch_destroy(pa_batch *b);

/* Launch shape of batched chain launches (no reference counterpart: the
 * reference runs one element per OpenMP thread): the block size (64, 128 or
 * 256 threads, or 0 for the default chosen from the batch's mean pattern
 * count) and the pattern-loop shape 0..PA_N_SHAPES-1, a pair (U nodes
 * pruned per round, K nodes per ancestral-sampling chunk): FP32 (2, 4),
 * (1, 4), (2, 8); FP64 (1, 2), (2, 2), (1, 4).  Neither changes any result
 * (each chain is one workgroup's, and the shapes only regroup independent
 * loads).  round_nodes / chunk_nodes report U and K of the shape at the
 * precision the autotuner ran; rate: the element-MCMC iterations/s it
 * measured. */
enum { PA_N_SHAPES = 3 };
typedef struct {
    int32_t block_threads, shape;
    int32_t round_nodes, chunk_nodes;
    double rate;
} pa_tuning;

pa_status pa_set_tuning(pa_engine *e, const pa_tuning *t);
pa_status pa_get_tuning(pa_engine *e, pa_tuning *t);
/* Times every block size (64, 128, 256) and shape on the batch (or, with
 * sample > 0, on `sample` of its elements evenly spaced by pattern count)
 * with chains of `states` states (0: 8), times the three fastest again twice
 * at 2 * states and keeps the best of those (rate: its iterations per second
 * there) for later pa_run_chains / pa_batch_create calls with
 * block_threads = 0. */
pa_status pa_autotune(pa_engine *e, const int32_t *ids, int32_t n, int32_t precision, int32_t sample,
                      int32_t states, pa_tuning *out);

/* Per-model chain summaries of one element from its traces (T = BURNIN+MCMC
 * states; Z is [T*N], max_Z is [N]).  vb: marginal log-likelihood of
 * BPP_C::Eval2 (src/PhyloAcc-ST/bpp_c2.cpp:14-353), whose fallback branch
 * re-evaluates P(X | Z, r*) on the device at `precision`.  medians[5] (n, c,
 * g, l, l2 rates) and postZ[N*4] (frequency of Z = -1, 0, 1, 2 after burn-in)
 * are BPP_C::Output_init's row (bpp_c2.cpp:739-785).  model: 0 = M0, 1 = M1,
 * 2 = M2.  Any output pointer may be NULL. */
pa_status pa_chain_summary(pa_engine *e, int32_t id, int32_t model, const double *full, const double *nrate,
                           const double *crate, const double *grate, const double *lrate, const double *l2rate,
                           const int8_t *Z, const int8_t *max_Z, int32_t precision, double *vb, double *medians,
                           double *postZ);

/* ------------------------------------------------------------------ serving
 * Continuous request loop over element chains.  The reference has no serving
 * mode: its OpenMP loop (src/PhyloAcc-ST/main.cpp:452-500) runs a fixed id
 * list to completion.  Here requests (already encoded element ids) arrive at
 * any time and wait in arrival order.  Up to `lanes` launches are in flight
 * at once, each on its own stream and holding a range of the server's HBM
 * pool (hbm_budget_bytes) while it runs; a free lane takes the waiting
 * requests that fit the largest free range, up to max_chains.  A lane is
 * started when nothing is in flight, or when fewer chains are in flight than
 * the chain kernel has resident workgroups on the device (retiring chains
 * have freed slots).  Every element retires on its own: the kernel writes its
 * three result rows into host-visible memory and then a completion flag, and
 * each pa_serve_step collects the elements whose flag is set.  A chain's
 * result does not depend on the lane or on what else is in flight: its
 * random streams are keyed by SEED and the element, one workgroup runs it,
 * and the block size is fixed per server. */
typedef struct pa_server pa_server;

typedef struct {
    int32_t precision;          /* PA_FP64 / PA_FP32 */
    int32_t block_threads;      /* 64, 128 or 256; 0 = the engine's tuned block size, else one chosen from the mean pattern count */
    int32_t lanes;              /* launches in flight at once, 1..PA_SERVE_MAX_LANES; 0 = PA_SERVE_DEFAULT_LANES */
    int32_t max_chains;         /* chains in flight at most; 0 = 4096 */
    double hbm_budget_bytes;    /* scratch pool shared by the lanes' launches; 0 = 4 GiB */
} pa_serve_config;

enum { PA_SERVE_MAX_LANES = 8, PA_SERVE_DEFAULT_LANES = 3 };

enum { PA_SERVE_CHAINS = 0, PA_SERVE_LOGLIK = 1 };

typedef struct {
    int64_t ticket;             /* from pa_serve_submit / pa_serve_submit_lengths */
    int32_t element;            /* element id */
    int32_t status;             /* PA_ELEM_OK / PA_ELEM_NUMERIC_FAIL */
    int32_t kind;               /* PA_SERVE_CHAINS: result[]; PA_SERVE_LOGLIK: loglik */
    int32_t pad;
    pa_chain_result result[3];  /* M0, M1, M2 */
    double loglik;
    double arrival_s, admit_s, finish_s;   /* pa_serve_clock() times */
} pa_serve_result;

typedef struct {
    int64_t submitted, admitted, finished, launches;
    int32_t waiting, in_flight;            /* in_flight: admitted, not yet retired */
    double scratch_bytes_in_use, device_ms;
    int32_t lanes_busy, pad;
} pa_serve_stats;

pa_status pa_serve_create(pa_engine *e, const pa_serve_config *cfg, pa_server **out);
/* Queue n element ids (FIFO).  arrival_s: per-request arrival times on the
 * pa_serve_clock() scale, or NULL for "now".  tickets: [n] or NULL. */
pa_status pa_serve_submit(pa_server *s, const int32_t *ids, int32_t n, const double *arrival_s, int64_t *tickets);
/* Log-likelihood requests for elements that carry their own tree (gene
 * trees, pa_elements_add_trees) at per-request branch lengths [n*N]: each
 * step prunes the admitted ones in one batched launch after the chains. */
pa_status pa_serve_submit_lengths(pa_server *s, const int32_t *ids, int32_t n, const double *lengths,
                                  const double *arrival_s, int64_t *tickets);
/* One turn of the request loop: collects the elements that finished, then
 * starts lanes for waiting requests; with wait = 0 it returns at once, with
 * wait = 1 it keeps turning until nothing is in flight (every admissible
 * request has finished). */
pa_status pa_serve_step(pa_server *s, int32_t wait, int32_t *n_finished);
/* Moves up to max finished results out of the server, oldest first. */
pa_status pa_serve_poll(pa_server *s, pa_serve_result *out, int32_t max, int32_t *n);
pa_status pa_serve_get_stats(pa_server *s, pa_serve_stats *out);
/* Waits for the launches in flight, then frees the server. */
void pa_serve_destroy(pa_server *s);
/* Monotonic host clock in seconds. */
double pa_serve_clock(void);

/* Number of visible HIP devices (0 when none); no reference counterpart. */
int32_t pa_device_count(void);

/* Host-side Philox4x32-10 uniform for stream slot (c0..c3) word `lane`, the
 * generator every device sampler draws from (no reference counterpart: the
 * reference draws from GSL's mt19937, src/PhyloAcc-ST/bpp_c.hpp:146-147). */
double pa_rng_uniform(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, int32_t lane);

#ifdef __cplusplus
}
#endif
#endif
