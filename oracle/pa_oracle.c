/*
 * pa_oracle.c -- CPU restatement of PhyloAcc-ST's per-element MCMC path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  Nothing in
 * phyloacc_amd/ links or calls this file.
 *
 * Everything is FP64, log space and one alignment site at a time, as in the
 * reference.  Each function cites the reference lines (src/PhyloAcc-ST/...)
 * it restates.  Departures, all documented where they occur:
 *   - P(t) = expm(Q t) is computed by scaling-and-squaring Taylor series
 *     rather than Armadillo's eig_gen (same function, independent method);
 *     entries are clamped at 0 before the log.
 *   - Random numbers: counter-based Philox4x32-10 with a fixed slot layout
 *     instead of GSL's mt19937 stream; samplers: categorical by inverse CDF,
 *     Gamma by Marsaglia-Tsang (2000) with Box-Muller normals, Beta as a
 *     ratio of Gammas.  Gamma log-densities are closed forms in log space.
 */
#include "pa_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <ctype.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NEG_INF (-INFINITY)

/* ------------------------------------------------------------------ RNG -- */

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    /* Salmon, Moraes, Dror, Shaw (SC'11), Philox4x32 with 10 rounds. */
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u01(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

static inline uint32_t tagword(int model, int tag) { return ((uint32_t)model << 24) | ((uint32_t)tag << 16); }

double or_uniform(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, int lane)
{
    uint32_t ctr[4] = {c0, c1, c2, c3}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    or_philox4x32_10(ctr, key, o);
    return u01(o[lane & 3]);
}

/* uniform for a fixed slot t of sub-stream c1 */
static inline double slot_u(uint64_t seed, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t t)
{
    return or_uniform(seed, t >> 2, c1, c2, c3, (int)(t & 3));
}

typedef struct { uint64_t seed; uint32_t c1, c2, c3, idx; } stream_t;

static inline double st_next(stream_t *s) { double u = slot_u(s->seed, s->c1, s->c2, s->c3, s->idx); s->idx++; return u; }

#define OR_GAMMA_MAXIT 64
/* Marsaglia & Tsang (2000), shape a > 0, scale b; for a < 1 the boost
 * U^(1/a) uses draw 0 and the main loop starts at draw 1. */
static double gamma_draw_c(stream_t *s, double a, double b, int *capped);
static double gamma_draw(stream_t *s, double a, double b) { return gamma_draw_c(s, a, b, NULL); }
/* capped (may be NULL): set to 1 when the rejection loop stops at its cap */
static double gamma_draw_c(stream_t *s, double a, double b, int *capped)
{
    double boost = 1.0;
    if (a < 1.0) { double u = st_next(s); boost = pow(u, 1.0 / a); a += 1.0; }
    double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (int it = 0; it < OR_GAMMA_MAXIT; it++) {
        double u1 = st_next(s), u2 = st_next(s);
        double x = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
        double v = 1.0 + c * x;
        if (v <= 0.0) continue;
        v = v * v * v;
        double u = st_next(s);
        if (u < 1.0 - 0.0331 * (x * x) * (x * x)) return d * v * b * boost;
        if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v * b * boost;
    }
    if (capped) *capped = 1;
    return d * b * boost;
}

double or_gamma_stream(uint64_t seed, uint32_t c1, uint32_t c2, uint32_t c3, double shape, double scale)
{
    stream_t s = {seed, c1, c2, c3, 0};
    return gamma_draw(&s, shape, scale);
}

static double beta_draw_c(uint64_t seed, uint32_t c1x, uint32_t c1y, uint32_t c2, uint32_t c3, double a, double b,
                          int *capped)
{
    stream_t sx = {seed, c1x, c2, c3, 0}, sy = {seed, c1y, c2, c3, 0};
    double x = gamma_draw_c(&sx, a, 1.0, capped), y = gamma_draw_c(&sy, b, 1.0, capped);
    return x / (x + y);
}

double or_beta_stream(uint64_t seed, uint32_t c1x, uint32_t c1y, uint32_t c2, uint32_t c3, double a, double b)
{
    stream_t sx = {seed, c1x, c2, c3, 0}, sy = {seed, c1y, c2, c3, 0};
    double x = gamma_draw(&sx, a, 1.0), y = gamma_draw(&sy, b, 1.0);
    return x / (x + y);
}

/* categorical over k unnormalised non-negative weights, inverse CDF */
static int categorical(const double *p, int k, double u)
{
    double tot = 0; int last = 0;
    for (int i = 0; i < k; i++) { tot += p[i]; if (p[i] > 0) last = i; }
    double target = u * tot, cum = 0;
    for (int i = 0; i < k - 1; i++) { cum += p[i]; if (target < cum) return i; }
    return last;
}

/* log of gsl_ran_gamma_pdf(x; a, scale b), closed form */
static double log_gamma_pdf(double x, double a, double b)
{
    if (x < 0) return NEG_INF;
    if (x == 0) { if (a == 1) return -log(b); return a > 1 ? NEG_INF : INFINITY; }
    return (a - 1) * log(x) - x / b - lgamma(a) - a * log(b);
}

static double lnbeta(double a, double b) { return lgamma(a) + lgamma(b) - lgamma(a + b); }

/* ---------------------------------------------------------- .mod text -- */
/* Restates LoadPhyloTree (newick.cpp:431-517), parseTree (newick.cpp:145-331)
 * and the TravelTree2/3 numbering (newick.cpp:382-428): post-order; leaves
 * 0..S-1 in encounter order, internal nodes S.. in post-order, root N-1;
 * children ordered by id (bpp.cpp:105-117); branch lengths rounded to float. */

typedef struct pnode { int leaf; int kids[2]; int nk; double dist; char name[64]; int id; } pnode;

typedef struct { const char *s; int pos; pnode *nodes; int n, cap; int err; } pstate;

static int pnew(pstate *p)
{
    if (p->n >= p->cap) { p->err = 1; return 0; }
    memset(&p->nodes[p->n], 0, sizeof(pnode));
    return p->n++;
}

/* parse one subtree in s[a..b) */
static int parse_sub(pstate *p, int a, int b)
{
    const char *s = p->s;
    int me = pnew(p); if (p->err) return 0;
    if (s[a] != '(') {
        int colon = -1;
        for (int i = a; i < b; i++) if (s[i] == ':') colon = i;
        int e = colon < 0 ? b : colon;
        int L = e - a; if (L > 63) L = 63;
        memcpy(p->nodes[me].name, s + a, L); p->nodes[me].name[L] = 0;
        p->nodes[me].dist = colon < 0 ? 0.0 : (double)(float)atof(s + colon + 1);
        p->nodes[me].leaf = 1;
        return me;
    }
    int depth = 0, start = a + 1, i;
    for (i = a; i < b; i++) {
        char c = s[i];
        if (c == '(') depth++;
        else if (c == ')') { depth--; if (depth == 0) break; }
        if ((c == ',' && depth == 1)) {
            if (p->nodes[me].nk >= 2) { p->err = 2; return me; }
            int k = parse_sub(p, start, i); p->nodes[me].kids[p->nodes[me].nk++] = k; start = i + 1;
        }
    }
    if (i >= b) { p->err = 3; return me; }
    if (p->nodes[me].nk >= 2) { p->err = 2; return me; }
    { int k = parse_sub(p, start, i); p->nodes[me].kids[p->nodes[me].nk++] = k; }
    if (p->nodes[me].nk != 2) { p->err = 2; return me; }
    /* label and distance after ')' */
    int j = i + 1, colon = -1;
    for (int q = j; q < b && s[q] != ';'; q++) if (s[q] == ':') { colon = q; break; }
    int e = colon < 0 ? j : colon;
    if (colon < 0) { e = j; while (e < b && s[e] != ';') e++; }
    int L = e - j; if (L > 63) L = 63;
    memcpy(p->nodes[me].name, s + j, L); p->nodes[me].name[L] = 0;
    if (colon >= 0) {
        p->nodes[me].dist = (double)(float)atof(s + colon + 1);
    }
    return me;
}

static void number_post(pstate *p, int v, int *leaf_id, int *int_id, int S)
{
    pnode *n = &p->nodes[v];
    for (int k = 0; k < n->nk; k++) number_post(p, n->kids[k], leaf_id, int_id, S);
    if (n->leaf) n->id = (*leaf_id)++;
    else n->id = S + (*int_id)++;
}

static void trim(char *t)
{
    int L = (int)strlen(t), a = 0;
    while (a < L && (t[a] == ' ' || t[a] == '\t' || t[a] == '\n')) a++;
    while (L > a && (t[L - 1] == ' ' || t[L - 1] == '\t' || t[L - 1] == '\n')) L--;
    memmove(t, t + a, L - a); t[L - a] = 0;
}

int or_parse_mod(const char *text, or_model *m, char *names, int max_names)
{
    memset(m, 0, sizeof(*m));
    char *buf = strdup(text); if (!buf) return -1;
    char *lines[4096]; int nl = 0;
    for (char *q = buf; nl < 4096;) {
        lines[nl++] = q;
        char *e = strchr(q, '\n');
        if (!e) break;
        *e = 0; q = e + 1;
    }
    char *tree = NULL;
    for (int li = 0; li < nl; li++) {
        if (lines[li][0] == 0) break;           /* newick.cpp:457 stops at the first empty line */
        trim(lines[li]);
        char first[64] = {0}; sscanf(lines[li], "%63s", first);
        if (lines[li][0] == ' ' || lines[li][0] == 0) continue;
        if (!strcmp(first, "BACKGROUND:")) {
            char *q = lines[li] + strlen("BACKGROUND:"); int ind = 0;
            while (ind < 4) { char *e; double v = strtod(q, &e); if (e == q) break; m->pi[ind++] = v; q = e; }
        } else if (!strcmp(first, "RATE_MAT:")) {
            for (int i = 0; i < 4 && li + 1 < nl; i++) {
                li++; char *q = lines[li]; int ind = 0;
                while (ind < 4) { char *e; double v = strtod(q, &e); if (e == q) break; m->Q[i * 4 + ind++] = v; q = e; }
            }
        } else if (!strcmp(first, "TREE:")) {
            char *q = lines[li] + 5; while (*q == ' ') q++;
            char *e = q; while (*e && *e != ' ') e++; *e = 0;
            tree = q;
        }
    }
    if (!tree) { free(buf); return -2; }
    int L = (int)strlen(tree);
    pstate p = {tree, 0, (pnode *)calloc(2 * OR_MAXN, sizeof(pnode)), 0, 2 * OR_MAXN, 0};
    int end = L; while (end > 0 && tree[end - 1] == ';') end--;
    int root = parse_sub(&p, 0, end);
    if (p.err) { free(p.nodes); free(buf); return -3; }
    int S = 0; for (int i = 0; i < p.n; i++) S += p.nodes[i].leaf;
    int N = 2 * S - 1;
    if (N > OR_MAXN || p.n != N) { free(p.nodes); free(buf); return -4; }
    int lid = 0, iid = 0; number_post(&p, root, &lid, &iid, S);
    m->S = S; m->N = N;
    for (int i = 0; i < N; i++) { m->children[i][0] = m->children[i][1] = -1; m->parent[i] = -1; }
    for (int v = 0; v < p.n; v++) {
        pnode *n = &p.nodes[v]; int id = n->id;
        m->dist[id] = n->dist;
        if (names && id < max_names) { strncpy(names + 64 * id, n->name, 63); names[64 * id + 63] = 0; }
        if (!n->leaf) {
            int a = p.nodes[n->kids[0]].id, b = p.nodes[n->kids[1]].id;
            if (a > b) { int t = a; a = b; b = t; }
            m->children[id][0] = a; m->children[id][1] = b;
            m->parent[a] = id; m->parent[b] = id;
        }
    }
    free(p.nodes); free(buf);
    return 0;
}

/* BPP::getUppertree (bpp.cpp:486-504): the species themselves, their
 * ancestors, and the root. */
void or_set_groups(or_model *m, const int *outgroup, int n_out, const int *conserve, int n_cons)
{
    int root = m->N - 1;
    memset(m->upper, 0, sizeof(m->upper)); memset(m->upper_conserve, 0, sizeof(m->upper_conserve));
    for (int i = 0; i < n_out; i++) for (int p = outgroup[i]; p != root && p >= 0; p = m->parent[p]) m->upper[p] = 1;
    m->upper[root] = 1;
    for (int i = 0; i < n_cons; i++) for (int p = conserve[i]; p != root && p >= 0; p = m->parent[p]) m->upper_conserve[p] = 1;
    m->upper_conserve[root] = 1;
    m->n_conserve = n_cons;
    for (int i = 0; i < n_cons; i++) m->conserve[i] = conserve[i];
}

/* ------------------------------------------------ transition matrices -- */

static void mat4_mul(const double *A, const double *B, double *C)
{
    double T[16];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) {
        double s = 0; for (int k = 0; k < 4; k++) s += A[i * 4 + k] * B[k * 4 + j];
        T[i * 4 + j] = s;
    }
    memcpy(C, T, sizeof T);
}

/* expm(Q t) by scaling and squaring with a degree-24 Taylor sum */
void or_expm(const double *Q, double t, double *P)
{
    double A[16], nrm = 0;
    for (int i = 0; i < 16; i++) A[i] = Q[i] * t;
    for (int i = 0; i < 4; i++) { double r = 0; for (int j = 0; j < 4; j++) r += fabs(A[i * 4 + j]); if (r > nrm) nrm = r; }
    int sq = 0;
    while (nrm > 0.125) { nrm *= 0.5; sq++; }
    double sc = ldexp(1.0, -sq);
    for (int i = 0; i < 16; i++) A[i] *= sc;
    double term[16], sum[16];
    for (int i = 0; i < 16; i++) { term[i] = (i % 5 == 0) ? 1.0 : 0.0; sum[i] = term[i]; }
    for (int k = 1; k <= 24; k++) {
        mat4_mul(term, A, term);
        for (int i = 0; i < 16; i++) { term[i] /= k; sum[i] += term[i]; }
    }
    for (int s = 0; s < sq; s++) mat4_mul(sum, sum, sum);
    memcpy(P, sum, sizeof sum);
}

/* log((V e^{L t} V^-1)^T), bpp.hpp:370-397 and bpp_c.cpp:187-228:
 * logPt[i*4+j] = log P_{j->i}(t).  Used for t > threshold only by callers. */
static void log_trans_q(const double *Q, double t, double *logPt)
{
    double P[16];
    or_expm(Q, t, P);
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) {
        double v = P[j * 4 + i]; if (v < 0) v = 0;
        logPt[i * 4 + j] = log(v);
    }
}

void or_log_trans(const or_model *m, double t, double *logPt)
{
    double P[16];
    or_expm(m->Q, t, P);
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) {
        double v = P[j * 4 + i]; if (v < 0) v = 0;
        logPt[i * 4 + j] = log(v);
    }
}

static void log_identity(double *L)
{
    for (int i = 0; i < 16; i++) L[i] = (i % 5 == 0) ? 0.0 : NEG_INF;
}

/* ------------------------------------------------------ log helpers -- */

/* BPP::log_multi (bpp.hpp:427-445): out_j = log sum_i exp(L(i,j) + y_i),
 * one global max over the 4x4 */
static void log_multi(const double *L, const double *y, double *out)
{
    double x[16], mx = NEG_INF;
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) {
        x[i * 4 + j] = L[i * 4 + j] + y[i];
        if (x[i * 4 + j] > mx) mx = x[i * 4 + j];
    }
    if (mx == NEG_INF) { for (int j = 0; j < 4; j++) out[j] = NEG_INF; return; }
    for (int j = 0; j < 4; j++) {
        double s = 0; for (int i = 0; i < 4; i++) s += exp(x[i * 4 + j] - mx);
        out[j] = log(s) + mx;
    }
}

/* BPP::log_multi2 (bpp.hpp:447-460): per-column max, 3x3 */
static void log_multi2(const double *L, const double *y, double *out)
{
    for (int j = 0; j < 3; j++) {
        double x[3], mx = NEG_INF;
        for (int i = 0; i < 3; i++) { x[i] = L[i * 3 + j] + y[i]; if (x[i] > mx) mx = x[i]; }
        double s = 0; for (int i = 0; i < 3; i++) s += exp(x[i] - mx);
        out[j] = log(s) + mx;
    }
}

/* BPP::log_exp_sum (bpp.hpp:471-479) */
static double log_exp_sum(const double *y, int k)
{
    double mx = NEG_INF;
    for (int i = 0; i < k; i++) if (y[i] > mx) mx = y[i];
    if (mx == NEG_INF) return NEG_INF;
    double s = 0; for (int i = 0; i < k; i++) s += exp(y[i] - mx);
    return log(s) + mx;
}

/* ----------------------------------------------------------- element -- */

struct or_element {
    int filtered, GG, S, N, root;
    int *Tg;                 /* [GG*N] */
    double *lam;             /* [GG*N*4] */
    unsigned char *code;     /* [GG*S] */
    int missing[OR_MAXN];
    int ch2[OR_MAXN][2], par2[OR_MAXN];
    double dist2[OR_MAXN];
    int nodes[OR_MAXN], n_nodes, inodes[OR_MAXN], n_inodes;
    int upper_c[OR_MAXN], upper_cons_c[OR_MAXN];   /* flags */
    double log_emission[OR_MAXN][3];
    /* chain state that persists across the three models of one element */
    int Z[OR_MAXN], fixZ[OR_MAXN];
    double TM[OR_MAXN][9];   /* log_TM_Int, [i*3+j] = log P(child i | parent j) */
    double Lnull[OR_MAXN][16], Lcons[OR_MAXN][16], Lneut[OR_MAXN][16];
    double *amb_null;        /* [GG*S*4] */
    int accept_n, accept_c;
    double prop_n, prop_c;
};

static void encode_char(int y, int gapchar, double *lam, int *Tg, unsigned char *code)
{
    /* bpp_c.hpp:213-261 */
    for (int b = 0; b < 4; b++) lam[b] = NEG_INF;
    switch (y) {
    case 'a': lam[0] = 0; *Tg = 0; *code = OR_A; break;
    case 'c': lam[1] = 0; *Tg = 1; *code = OR_C; break;
    case 'g': lam[2] = 0; *Tg = 2; *code = OR_G; break;
    case 't': lam[3] = 0; *Tg = 3; *code = OR_T; break;
    case 'r': lam[0] = 0; lam[2] = 0; *Tg = -1; *code = OR_R; break;
    case 'y': lam[1] = 0; lam[3] = 0; *Tg = -1; *code = OR_Y; break;
    case 'k': lam[2] = 0; lam[3] = 0; *Tg = -1; *code = OR_K; break;
    case 'm': lam[0] = 0; lam[1] = 0; *Tg = -1; *code = OR_M; break;
    case 's': lam[1] = 0; lam[2] = 0; *Tg = -1; *code = OR_S; break;
    case 'w': lam[0] = 0; lam[3] = 0; *Tg = -1; *code = OR_W; break;
    default:
        for (int b = 0; b < 4; b++) lam[b] = 0;
        if (y == gapchar) { *Tg = 4; *code = OR_GAP; } else { *Tg = 5; *code = OR_OTHER; }
    }
}

/* BPP_C::getSubtree_missing (bpp_c.cpp:64-113) */
static void subtree_missing(or_element *e, int ss, const int *upper, int child)
{
    if (!upper[ss] || ss < e->S) return;
    int s = ss;
    while (upper[s] && s >= e->S) {
        int p0 = e->ch2[s][0], p1 = e->ch2[s][1];
        if (e->missing[p0] && !e->missing[p1]) {
            e->par2[p0] = -1;
            e->dist2[p1] += e->dist2[s];
            if (s == e->root) e->root = p1;
            else { e->ch2[e->par2[ss]][child] = p1; e->par2[p1] = e->par2[ss]; }
            s = p1;
        } else if (!e->missing[p0] && e->missing[p1]) {
            e->par2[p1] = -1;
            e->dist2[p0] += e->dist2[s];
            if (s == e->root) e->root = p0;
            else { e->ch2[e->par2[ss]][child] = p0; e->par2[p0] = e->par2[ss]; }
            s = p0;
        } else {
            subtree_missing(e, p0, upper, 0);
            subtree_missing(e, p1, upper, 1);
            return;
        }
    }
}

/* BPP_C::getSubtree (bpp_c.cpp:44-60): post-order via children2 */
static void post_order(or_element *e, int j)
{
    for (int c = 0; c < 2; c++) if (e->ch2[j][c] != -1) post_order(e, e->ch2[j][c]);
    e->nodes[e->n_nodes++] = j;
}

/* BPP_C constructor (bpp_c.hpp:143-476) for element rows[s][start..start+GG) */
or_element *or_encode(const or_model *m, const or_options *o, const char *const *rows, long start, int GG)
{
    or_element *e = (or_element *)calloc(1, sizeof(or_element));
    int S = m->S, N = m->N;
    e->S = S; e->N = N; e->GG = GG; e->root = N - 1;
    if (GG < o->min_len) { e->filtered = 1; return e; }
    int *Tg = (int *)malloc(sizeof(int) * (size_t)GG * N);
    double *lam = (double *)malloc(sizeof(double) * (size_t)GG * N * 4);
    unsigned char *code = (unsigned char *)malloc((size_t)GG * S);
    for (size_t i = 0; i < (size_t)GG * N; i++) Tg[i] = -1;
    for (int s = 0; s < S; s++) for (int g = 0; g < GG; g++) {
        int y = tolower((unsigned char)rows[s][start + g]);
        encode_char(y, o->gapchar, lam + ((size_t)g * N + s) * 4, Tg + (size_t)g * N + s, code + (size_t)g * S + s);
    }
    /* trimming, bpp_c.hpp:268-301 */
    if (o->trim_gap_percent < 1) {
        int keep = 0;
        char *drop = (char *)calloc(GG, 1);
        for (int g = 0; g < GG; g++) {
            int mis = 0;
            for (int s = 0; s < S; s++) if (Tg[(size_t)g * N + s] >= 4) mis++;
            if (mis > S * o->trim_gap_percent) drop[g] = 1; else keep++;
        }
        if (keep < o->min_len) { free(drop); free(Tg); free(lam); free(code); e->filtered = 1; return e; }
        int w = 0;
        for (int g = 0; g < GG; g++) {
            if (drop[g]) continue;
            if (w != g) {
                memmove(Tg + (size_t)w * N, Tg + (size_t)g * N, sizeof(int) * N);
                memmove(lam + (size_t)w * N * 4, lam + (size_t)g * N * 4, sizeof(double) * N * 4);
                memmove(code + (size_t)w * S, code + (size_t)g * S, S);
            }
            w++;
        }
        free(drop);
        GG = keep; e->GG = GG;
    }
    /* gap-only counts and missing flags, bpp_c.hpp:304-351 */
    int num_missing[OR_MAXN] = {0};
    for (int s = 0; s < S; s++) for (int g = 0; g < GG; g++) if (Tg[(size_t)g * N + s] == 4) num_missing[s]++;
    for (int i = 0; i < N; i++) {
        e->ch2[i][0] = m->children[i][0]; e->ch2[i][1] = m->children[i][1];
        e->par2[i] = m->parent[i]; e->dist2[i] = m->dist[i];
    }
    for (int s = S; s < N; s++) {
        for (int c = 0; c < 2; c++) {
            int chi = e->ch2[s][c];
            if (chi < S && num_missing[chi] > o->gap_prop * GG) e->missing[chi] = 1;
        }
        if (e->missing[e->ch2[s][0]] && e->missing[e->ch2[s][1]]) e->missing[s] = 1;
    }
    /* CONSERVE filter, bpp_c.hpp:361-373 */
    int ct = 0;
    for (int i = 0; i < m->n_conserve; i++) if (e->missing[m->conserve[i]]) ct++;
    if (ct > o->conserve_prop * m->n_conserve) { free(Tg); free(lam); free(code); e->filtered = 1; return e; }
    /* collapse, bpp_c.hpp:379-390 */
    if (o->prune_tree) { int all[OR_MAXN]; for (int i = 0; i < N; i++) all[i] = 1; subtree_missing(e, N - 1, all, -1); }
    else subtree_missing(e, N - 1, m->upper, -1);
    e->par2[e->root] = N;
    post_order(e, e->root);
    for (int i = 0; i < e->n_nodes; i++) {
        int s = e->nodes[i];
        if (s >= S) e->inodes[e->n_inodes++] = s;
        e->upper_c[s] = m->upper[s];
        e->upper_cons_c[s] = m->upper_conserve[s];
    }
    /* internal nodes whose children are both >= num_base, bpp_c.hpp:417-428 */
    for (int k = 0; k < e->n_inodes; k++) {
        int s = e->inodes[k], a = e->ch2[s][0], b = e->ch2[s][1];
        for (int g = 0; g < GG; g++)
            if (Tg[(size_t)g * N + a] >= 4 && Tg[(size_t)g * N + b] >= 4) Tg[(size_t)g * N + s] = 4;
    }
    /* missing-node emissions, bpp_c.hpp:451-474 */
    for (int s = 0; s < N; s++) {
        e->log_emission[s][0] = e->log_emission[s][1] = e->log_emission[s][2] = 0;
        if (e->missing[s]) {
            e->log_emission[s][2] = log(o->nconstomis);
            e->log_emission[s][1] = log(o->constomis);
            e->log_emission[s][0] = log(o->nconstomis);
        }
    }
    for (int s = 0; s < N; s++) for (int k = 0; k < 9; k++) e->TM[s][k] = 0;   /* zeros<mat>(3,3) */
    e->Tg = Tg; e->lam = lam; e->code = code;
    e->amb_null = (double *)calloc((size_t)GG * S * 4, sizeof(double));
    return e;
}

void or_free(or_element *e)
{
    if (!e) return;
    free(e->Tg); free(e->lam); free(e->code); free(e->amb_null); free(e);
}

int or_elem_filtered(const or_element *e) { return e->filtered; }

int or_elem_info(const or_element *e, int *out)
{
    out[0] = e->GG; out[1] = e->root; out[2] = e->n_nodes; out[3] = e->n_inodes;
    out[4] = e->S; out[5] = e->N; out[6] = e->filtered; out[7] = 0;
    return 0;
}

void or_elem_arrays(const or_element *e, int *children2, int *parent2, double *dist2,
                    int *missing, int *nodes, int *internal, int *Tg, unsigned char *codes)
{
    int N = e->N;
    for (int i = 0; i < N; i++) {
        if (children2) { children2[2 * i] = e->ch2[i][0]; children2[2 * i + 1] = e->ch2[i][1]; }
        if (parent2) parent2[i] = e->par2[i];
        if (dist2) dist2[i] = e->dist2[i];
        if (missing) missing[i] = e->missing[i];
    }
    if (nodes) for (int i = 0; i < e->n_nodes; i++) nodes[i] = e->nodes[i];
    if (internal) for (int i = 0; i < e->n_inodes; i++) internal[i] = e->inodes[i];
    if (!e->filtered) {
        if (Tg) memcpy(Tg, e->Tg, sizeof(int) * (size_t)e->GG * N);
        if (codes) memcpy(codes, e->code, (size_t)e->GG * e->S);
    }
}

#define LAM(e, g, s) ((e)->lam + (((size_t)(g) * (e)->N + (s)) * 4))
#define TG(e, g, s) ((e)->Tg[(size_t)(g) * (e)->N + (s)])

/* class of rate on branch above s: Z=2 -> neut (r2), Z=0 -> null (1), else cons (r1) */
static const double *Lclass(const or_element *e, int s)
{
    int z = e->Z[s];
    return z == 2 ? e->Lneut[s] : (z == 0 ? e->Lnull[s] : e->Lcons[s]);
}

/* BPP_C::Update_Tg(tosample=false) for one site, bpp_c.cpp:647-671 */
static void prune_site(or_element *e, int g, const char *visited, const double *log_pi)
{
    for (int k = 0; k < e->n_inodes; k++) {
        int s = e->inodes[k];
        if (visited[s] || e->missing[s] || TG(e, g, s) >= 4) continue;
        double *ls = LAM(e, g, s);
        for (int c = 0; c < 2; c++) {
            int chi = e->ch2[s][c];
            if (e->missing[chi] || TG(e, g, chi) >= 4) continue;
            double o[4]; log_multi(Lclass(e, chi), LAM(e, g, chi), o);
            for (int j = 0; j < 4; j++) ls[j] += o[j];
        }
    }
    if (!visited[e->root]) { double *lr = LAM(e, g, e->root); for (int j = 0; j < 4; j++) lr[j] += log_pi[j]; }
}

static void zero_internal(or_element *e)
{
    for (int g = 0; g < e->GG; g++) for (int s = e->S; s < e->N; s++) { double *l = LAM(e, g, s); l[0] = l[1] = l[2] = l[3] = 0; }
}

static void model_caches(const or_model *m, or_element *e, double nrate, double crate)
{
    /* BPP null cache (bpp.hpp:366-397) then BPP_C::initMCMC caches (bpp_c.cpp:180-228) */
    for (int s = 0; s < e->N; s++) {
        if (s < e->N - 1 && m->dist[s] > 1e-10) or_log_trans(m, m->dist[s], e->Lnull[s]);
        else if (s < e->N - 1) log_identity(e->Lnull[s]);
        else for (int i = 0; i < 16; i++) e->Lnull[s][i] = 0;
    }
    for (int k = 0; k < e->n_nodes - 1; k++) {
        int s = e->nodes[k];
        if (e->dist2[s] > 1e-10) {
            if (e->dist2[s] != m->dist[s]) or_log_trans(m, e->dist2[s], e->Lnull[s]);
            or_log_trans(m, e->dist2[s] * nrate, e->Lneut[s]);
            or_log_trans(m, e->dist2[s] * crate, e->Lcons[s]);
        } else {
            memcpy(e->Lcons[s], e->Lnull[s], sizeof(double) * 16);
            memcpy(e->Lneut[s], e->Lnull[s], sizeof(double) * 16);
        }
    }
}

double or_loglik(const or_model *m, or_element *e, const int *Z, double r1, double r2)
{
    /* log_f_Xz full form, bpp_c2.cpp:392-425, with caches at (r2, r1) */
    if (e->filtered) return NAN;
    model_caches(m, e, r2, r1);
    for (int s = 0; s < e->N; s++) e->Z[s] = Z[s];
    zero_internal(e);
    double logpi[4]; for (int j = 0; j < 4; j++) logpi[j] = log(m->pi[j]);
    char visited[OR_MAXN] = {0};
    double res = 0;
    for (int g = 0; g < e->GG; g++) {
        prune_site(e, g, visited, logpi);
        res += log_exp_sum(LAM(e, g, e->root), 4);
    }
    return res;
}

/* ------------------------------------------------- GT base frequencies -- */

/* BPP_C::getQmat (src/PhyloAcc-GT/bpp_c.cpp:160-185) at pi = (a, 1/2 - a,
 * 1/2 - a, a), with BPP's exchangeabilities from the model's RATE_MAT and
 * BACKGROUND (src/PhyloAcc-GT/bpp.hpp:346-353: inst(3) = inst(0), inst(4) =
 * Q(1,2) / pi(2), inst(5) = inst(1)). */
static void gt_model_at(const or_model *m, double a, or_model *out)
{
    double inst[6];
    inst[0] = m->Q[0 * 4 + 1] / m->pi[1];
    inst[1] = m->Q[0 * 4 + 2] / m->pi[2];
    inst[2] = m->Q[0 * 4 + 3] / m->pi[3];
    inst[3] = inst[0];
    inst[4] = m->Q[1 * 4 + 2] / m->pi[2];
    inst[5] = inst[1];
    *out = *m;
    double *pi = out->pi, *Q = out->Q;
    pi[0] = pi[3] = a;
    pi[1] = pi[2] = 0.5 - a;
    Q[0 * 4 + 0] = -(inst[0] * pi[1] + inst[1] * pi[2] + inst[2] * pi[3]);
    Q[0 * 4 + 1] = inst[0] * pi[1]; Q[0 * 4 + 2] = inst[1] * pi[2]; Q[0 * 4 + 3] = inst[2] * pi[3];
    Q[1 * 4 + 0] = inst[0] * pi[0];
    Q[1 * 4 + 1] = -(inst[0] * pi[0] + inst[3] * pi[2] + inst[4] * pi[3]);
    Q[1 * 4 + 2] = inst[3] * pi[2]; Q[1 * 4 + 3] = inst[4] * pi[3];
    Q[2 * 4 + 0] = inst[1] * pi[0]; Q[2 * 4 + 1] = inst[3] * pi[1];
    Q[2 * 4 + 2] = -(inst[1] * pi[0] + inst[3] * pi[1] + inst[5] * pi[3]);
    Q[2 * 4 + 3] = inst[5] * pi[3];
    Q[3 * 4 + 0] = inst[2] * pi[0]; Q[3 * 4 + 1] = inst[4] * pi[1]; Q[3 * 4 + 2] = inst[5] * pi[2];
    Q[3 * 4 + 3] = -(inst[2] * pi[0] + inst[4] * pi[1] + inst[5] * pi[2]);
}

double or_gt_loglik_pi(const or_model *m, double a, int S, const int *gchildren, int K, const int *seg_gene,
                       const int *seg_sp, const double *seg_len, const int *Z, double r1, double r2,
                       const char *const *rows, long start, int GG, int gapchar, double gap_prop)
{
    or_model mm;
    gt_model_at(m, a, &mm);
    return or_gt_loglik(&mm, S, gchildren, K, seg_gene, seg_sp, seg_len, Z, r1, r2, rows, start, GG, gapchar, gap_prop);
}

static double log_beta_pdf(double x, double a, double b)
{
    return lgamma(a + b) - lgamma(a) - lgamma(b) + (a - 1) * log(x) + (b - 1) * log1p(-x);
}

/* BPP_C::sample_pi, src/PhyloAcc-GT/bpp_c.cpp:1508-1597, for one element at
 * sweep `sweep` over its full length.  *piA and *loglik are the current a and
 * log-likelihood and are updated; returns 1 when accepted (an unchanged
 * proposal counts as accepted).  Draws: tag PI; the step's uniform on
 * sub-stream 0, the decision's on 1, the resample Beta's variates on 2, 3. */
int or_gt_sample_pi(const or_model *m, uint64_t seed, int sweep, double pi_delta, double prior_a, double prior_b,
                    int resample, double *piA, double *loglik, int S, const int *gchildren, int K,
                    const int *seg_gene, const int *seg_sp, const double *seg_len, const int *Z, double r1,
                    double r2, const char *const *rows, long start, int GG, int gapchar, double gap_prop)
{
    uint32_t c3 = tagword(0, OR_TAG_PI), c2 = (uint32_t)sweep;
    double a = *piA, t, log_p_prop = 0, log_p_cur = 0;
    if (resample) {
        /* obsACGT: a, c, g, t of the leaves over the element (Gibbs, bpp_c.cpp:388-394) */
        double cnt[4] = {0, 0, 0, 0};
        for (int s = 0; s < S; s++)
            for (int g = 0; g < GG; g++) {
                int ch = tolower((unsigned char)rows[s][start + g]);
                int b = ch == 'a' ? 0 : ch == 'c' ? 1 : ch == 'g' ? 2 : ch == 't' ? 3 : -1;
                if (b >= 0) cnt[b] += 1;
            }
        t = beta_draw_c(seed, 2, 3, c2, c3, prior_a + cnt[0] + cnt[3], prior_b + cnt[1] + cnt[2], NULL) / 2;
    } else {
        t = a + (2.0 * slot_u(seed, 0, c2, c3, 0) - 1) * pi_delta;
        log_p_prop = -log(2.0 * pi_delta);
        if (t < 0.05) { log_p_prop += log(0.05 - (a - pi_delta)); t = 0.05; }
        else if (t > 0.45) { log_p_prop += log((a + pi_delta) - 0.45); t = 0.45; }
        log_p_cur = -log(2.0 * pi_delta);
        if (a == 0.05 && (t - pi_delta) < 0.05) log_p_cur += log(0.05 - (t - pi_delta));
        else if (a == 0.45 && (t + pi_delta) > 0.45) log_p_cur += log(t + pi_delta - 0.45);
    }
    if (t == a) return 1;
    double lnew = or_gt_loglik_pi(m, t, S, gchildren, K, seg_gene, seg_sp, seg_len, Z, r1, r2, rows, start, GG,
                                  gapchar, gap_prop);
    if (resample) { *piA = t; *loglik = lnew; return 1; }
    double ratio = lnew - *loglik + log_beta_pdf(2 * t, prior_a, prior_b) - log_beta_pdf(2 * a, prior_a, prior_b)
                   + log_p_cur - log_p_prop;
    if (log(slot_u(seed, 1, c2, c3, 0)) < ratio) { *piA = t; *loglik = lnew; return 1; }
    return 0;
}

/* The Wang-Landau indicator step of PhyloAcc-GT's Gibbs (src/PhyloAcc-GT/
 * bpp_c.cpp:399-440) for one element at sweep m.  st[7]: log normalising
 * constants (2), momenta (2), learning rate, visit counts (2); set at m = 0
 * as bpp_c.cpp:423-427 does.  Returns the indicator (tag WL, sub-stream 0). */
int or_wl_step(uint64_t seed, int m, const double *loglik, const double *logpz, double *st)
{
    if (m == 0) {
        st[0] = loglik[0] + logpz[0]; st[1] = loglik[1] + logpz[1];
        st[2] = st[3] = 0; st[4] = 1; st[5] = st[6] = 0;
    }
    double y[2] = {loglik[0] + logpz[0] - st[0], loglik[1] + logpz[1] - st[1]}, prob[2] = {0, 0};
    double x = log_exp_sum(y, 2);                 /* BPP::log_sample_norm, GT bpp.hpp:497-503 */
    if (x != INFINITY) { prob[0] = exp(y[0] - x); prob[1] = exp(y[1] - x); }
    int ind = categorical(prob, 2, slot_u(seed, 0, (uint32_t)m, tagword(0, OR_TAG_WL), 0));
    const double beta = 0.9;
    st[2] = beta * st[2] - st[4] * prob[0];
    st[3] = beta * st[3] - st[4] * prob[1];
    st[5 + ind] += 1;
    st[0] -= st[2];
    st[1] -= st[3];
    double z = log_exp_sum(st, 2);
    st[0] -= z;
    st[1] -= z;
    double rr = st[5] / st[6];
    if (rr > 2.0 / 3.0 && rr < 3.0 / 2.0) {
        st[4] = 1 / (1 + 1 / st[4]);
        st[5] = st[6] = 0;
    }
    return ind;
}

/* --------------------------------------------------------------- gBGC -- */

/* BPP_C::getRate, src/PhyloAcc-ST-GBGC/SRC/bpp_c.cpp:13-29: B holds +b on
 * A->C, A->G, T->C, T->G and -b on C->A, C->T, G->A, G->T (bases A, C, G, T),
 * s is added to every entry, each entry x becomes x / (1 - e^-x) (1 at 0),
 * Q0 is scaled entrywise and the row sums are taken off the diagonal. */
void or_gbgc_rate(const double *Q0, double s, double b, double *Q)
{
    if (s == 0 && b == 0) { memcpy(Q, Q0, 16 * sizeof(double)); return; }
    double B[16] = {0};
    B[0 * 4 + 1] = b; B[0 * 4 + 2] = b; B[3 * 4 + 1] = b; B[3 * 4 + 2] = b;
    B[1 * 4 + 0] = -b; B[1 * 4 + 3] = -b; B[2 * 4 + 0] = -b; B[2 * 4 + 3] = -b;
    for (int i = 0; i < 16; i++) B[i] += s;
    for (int i = 0; i < 16; i++) B[i] = B[i] == 0 ? 1.0 : B[i] / (1 - exp(-B[i]));
    for (int i = 0; i < 16; i++) Q[i] = Q0[i] * B[i];
    for (int i = 0; i < 4; i++) {
        double c = 0;
        for (int j = 0; j < 4; j++) c += Q[i * 4 + j];
        Q[i * 4 + i] -= c;
    }
}

/* Pruning under the gBGC model: the six rate matrices of BPP_C::Gibbs
 * (src/PhyloAcc-ST-GBGC/SRC/bpp_c.cpp:191-197: background, conserved and
 * accelerated selection, each without and with the bias) and every
 * non-missing branch's log transition at its length (init_cache_TM,
 * bpp_c.cpp:1563-1621, computed for every length), then the same recursion
 * and site sum as or_loglik.  sel = {conserved selection, accelerated
 * selection, bias}; B[s] != 0 selects the biased class of branch s. */
double or_loglik_gbgc(const or_model *m, or_element *e, const int *Z, const int *B, const double *sel)
{
    if (e->filtered) return NAN;
    double Qk[6][16];
    for (int k = 0; k < 6; k++)
        or_gbgc_rate(m->Q, k % 3 == 0 ? 0.0 : (k % 3 == 1 ? sel[0] : sel[1]), k >= 3 ? sel[2] : 0.0, Qk[k]);
    for (int s = 0; s < e->N; s++) e->Z[s] = Z[s];
    for (int k = 0; k < e->n_nodes - 1; k++) {
        int s = e->nodes[k];
        if (e->missing[s]) continue;
        int cls = (Z[s] == 2 ? 2 : (Z[s] == 0 ? 0 : 1)) + (B[s] ? 3 : 0);
        double *dst = (double *)Lclass(e, s);   /* the table prune_site reads for branch s */
        log_trans_q(Qk[cls], e->dist2[s], dst);
    }
    zero_internal(e);
    double logpi[4]; for (int j = 0; j < 4; j++) logpi[j] = log(m->pi[j]);
    char visited[OR_MAXN] = {0};
    double res = 0;
    for (int g = 0; g < e->GG; g++) {
        prune_site(e, g, visited, logpi);
        res += log_exp_sum(LAM(e, g, e->root), 4);
    }
    return res;
}

/* ------------------------------------------------------------- chain -- */

typedef struct {
    const or_model *m; const or_options *o; or_element *e;
    int resZ, ms, T, m_cur;
    double logpi[4];
    double *tl, *tf, *tn, *tc, *tg, *tlr, *tl2; signed char *tZ;
    double MaxLoglik; int Max_m; int Max_Z[OR_MAXN];
    int flags;                             /* OR_FLAG_* bits (pa_oracle.h) */
} chain_t;

static void set_TM(or_element *e, int s, double g, double l, double l2, int init)
{
    double *T = e->TM[s];
    if (e->fixZ[s] == 1) {
        if (init) T[0 * 3 + 1] = NEG_INF;
        T[1 * 3 + 1] = 0; T[2 * 3 + 1] = NEG_INF;
        T[0 * 3 + 0] = log(1 - g); T[1 * 3 + 0] = log(g); T[2 * 3 + 0] = NEG_INF;
    } else {
        T[0 * 3 + 0] = log(1 - g); T[1 * 3 + 0] = log(g); T[2 * 3 + 0] = NEG_INF;
        double y = 1 - l;
        if (init) T[0 * 3 + 1] = NEG_INF;
        T[1 * 3 + 1] = log(y); T[2 * 3 + 1] = log(1 - y);
        if (init) T[0 * 3 + 2] = NEG_INF;
        T[1 * 3 + 2] = log(l2); T[2 * 3 + 2] = log(1 - l2);
    }
}

/* BPP_C::initMCMC (bpp_c.cpp:131-343), iter == 0 */
static void init_mcmc(chain_t *c)
{
    or_element *e = c->e; const or_options *o = c->o;
    e->prop_c = 1.2;
    e->prop_n = (double)1 / e->GG * 10;
    for (int s = 0; s < e->N; s++) e->Z[s] = -1;
    for (int k = 0; k < e->n_nodes; k++) e->Z[e->nodes[k]] = 1;
    e->Z[e->root] = 0;
    for (int s = 0; s < e->N; s++) e->fixZ[s] = -1;
    for (int k = 0; k < e->n_nodes; k++) if (e->upper_c[e->nodes[k]]) e->fixZ[e->nodes[k]] = 1;
    e->fixZ[e->root] = 1;
    c->MaxLoglik = NEG_INF; c->Max_m = 0;
    double lr2 = o->hyper_lrate2_a == 0 ? 0 : o->init_lrate2;
    c->tn[0] = o->init_nrate; c->tc[0] = o->init_crate;
    c->tlr[0] = o->init_lrate; c->tl2[0] = lr2; c->tg[0] = o->init_grate;
    model_caches(c->m, e, o->init_nrate, o->init_crate);
    zero_internal(e);
    /* getEmission_ambig, bpp_c.cpp:116-128 */
    for (int s = 0; s < e->S; s++) {
        if (e->missing[s]) continue;
        for (int g = 0; g < e->GG; g++) if (TG(e, g, s) == -1) log_multi(e->Lnull[s], LAM(e, g, s), e->amb_null + ((size_t)g * e->S + s) * 4);
    }
    if (c->resZ == 2) {
        for (int k = 0; k < e->n_nodes; k++) if (e->upper_cons_c[e->nodes[k]]) e->fixZ[e->nodes[k]] = 1;
    } else if (c->resZ == 0) {
        for (int k = 0; k < e->n_nodes - 1; k++) e->fixZ[e->nodes[k]] = 1;
    }
    for (int k = 0; k < e->n_nodes; k++) set_TM(e, e->nodes[k], c->tg[0], c->tlr[0], c->tl2[0], 1);
}

/* BPP_C::MonitorChain (bpp_c.cpp:741-834) */
static void monitor(chain_t *c, int m, double *loglik)
{
    or_element *e = c->e; const or_options *o = c->o;
    for (int s = 0; s < e->N; s++) c->tZ[(size_t)m * e->N + s] = (signed char)e->Z[s];
    if (m == 0) {
        double t = 0;
        for (int g = 0; g < e->GG; g++) t += log_exp_sum(LAM(e, g, e->root), 4);
        c->tl[0] = t; *loglik = t;
    } else c->tl[m] = *loglik;
    if (isnan(c->tl[m])) c->flags |= OR_FLAG_NAN;
    else if (isinf(c->tl[m])) c->flags |= c->tl[m] > 0 ? OR_FLAG_POS_INF : OR_FLAG_NEG_INF;
    c->tf[m] = c->tl[m] + log_gamma_pdf(c->tc[m], o->cprior_a, o->cprior_b);
    if (c->resZ != 0) c->tf[m] += log_gamma_pdf(c->tn[m], o->nprior_a, o->nprior_b);
    if (m >= o->burnin && c->tf[m] > c->MaxLoglik) {
        c->MaxLoglik = c->tf[m]; c->Max_m = m;
        for (int s = 0; s < e->N; s++) c->Max_Z[s] = e->Z[s];
    }
}

/* BPP_C::Update_Tg(tosample=true), bpp_c.cpp:672-733; the uniform for
 * site g and internal node at sampling position t (0 = root) is slot t of
 * sub-stream g in the ANC stream of sweep m. */
static void sample_site(chain_t *c, int g)
{
    or_element *e = c->e;
    uint32_t c3 = tagword(c->ms, OR_TAG_ANC), c2 = (uint32_t)c->m_cur;
    double p[4], *lr = LAM(e, g, e->root), mx = NEG_INF;
    for (int j = 0; j < 4; j++) if (lr[j] > mx) mx = lr[j];
    for (int j = 0; j < 4; j++) p[j] = exp(lr[j] - mx);
    TG(e, g, e->root) = categorical(p, 4, slot_u(c->o->seed, (uint32_t)g, c2, c3, 0));
    int t = 1;
    for (int k = e->n_inodes - 2; k >= 0; k--, t++) {
        int s = e->inodes[k], par = e->par2[s];
        if (e->missing[s] || TG(e, g, s) >= 4) continue;
        const double *L = Lclass(e, s);
        int tp = TG(e, g, par);
        double *ls = LAM(e, g, s), y[4];
        mx = NEG_INF;
        for (int i = 0; i < 4; i++) { y[i] = L[i * 4 + tp] + ls[i]; if (y[i] > mx) mx = y[i]; }
        for (int i = 0; i < 4; i++) p[i] = exp(y[i] - mx);
        TG(e, g, s) = categorical(p, 4, slot_u(c->o->seed, (uint32_t)g, c2, c3, (uint32_t)t));
    }
}

/* BPP_C::getEmission (bpp_c2.cpp:355-390) */
static void get_emission(chain_t *c)
{
    or_element *e = c->e;
    for (int k = 0; k < e->n_nodes - 1; k++) {
        int s = e->nodes[k], par = e->par2[s];
        if (e->missing[s]) continue;
        double e0 = 0, e1 = 0, e2 = 0;
        for (int g = 0; g < e->GG; g++) {
            int ts = TG(e, g, s), tp = TG(e, g, par);
            if (ts >= 4) continue;
            if (ts == -1) {
                double y[4];
                e0 += e->amb_null[((size_t)g * e->S + s) * 4 + tp];
                for (int i = 0; i < 4; i++) y[i] = e->Lcons[s][i * 4 + tp] + LAM(e, g, s)[i];
                e1 += log_exp_sum(y, 4);
                for (int i = 0; i < 4; i++) y[i] = e->Lneut[s][i * 4 + tp] + LAM(e, g, s)[i];
                e2 += log_exp_sum(y, 4);
            } else {
                e0 += e->Lnull[s][ts * 4 + tp];
                e1 += e->Lcons[s][ts * 4 + tp];
                e2 += e->Lneut[s][ts * 4 + tp];
            }
        }
        e->log_emission[s][0] = e0; e->log_emission[s][1] = e1; e->log_emission[s][2] = e2;
    }
}

/* BPP_C::Update_Z_subtree (bpp_c.cpp:1182-1293); the uniform for the node at
 * position t of the top-down order (0 = root) is slot t of the Z stream. */
static int update_Z_core(chain_t *c, int *changed);

static int update_Z(chain_t *c, int *changed)
{
    get_emission(c);
    return update_Z_core(c, changed);
}

/* the message passing and draws of Update_Z_subtree, given log_emission */
static int update_Z_core(chain_t *c, int *changed)
{
    or_element *e = c->e;
    double b[OR_MAXN][3];
    for (int k = 0; k < e->n_nodes; k++) {
        int s = e->nodes[k];
        for (int j = 0; j < 3; j++) b[s][j] = e->log_emission[s][j];
        if (e->fixZ[s] == 1) b[s][2] = NEG_INF;
        else if (e->fixZ[s] == 2) b[s][0] = NEG_INF;
    }
    for (int k = 0; k < e->n_nodes - 1; k++) {
        int s = e->nodes[k], p = e->par2[s];
        double o[3]; log_multi2(e->TM[s], b[s], o);
        for (int j = 0; j < 3; j++) b[p][j] += o[j];
    }
    double prior[3] = {log(0.5), log(0.5), NEG_INF};
    int r = e->root, nch = 0;
    for (int j = 0; j < 3; j++) b[r][j] += prior[j];
    uint32_t c3 = tagword(c->ms, OR_TAG_Z), c2 = (uint32_t)c->m_cur;
    double p[3], mx = NEG_INF;
    for (int j = 0; j < 3; j++) if (b[r][j] > mx) mx = b[r][j];
    for (int j = 0; j < 3; j++) p[j] = exp(b[r][j] - mx);
    int nn = categorical(p, 3, slot_u(c->o->seed, 0, c2, c3, 0));
    if (e->Z[r] != nn) changed[nch++] = r;
    e->Z[r] = nn;
    int t = 1;
    for (int k = e->n_nodes - 2; k >= 0; k--, t++) {
        int s = e->nodes[k], par = e->par2[s];
        double y[3]; mx = NEG_INF;
        for (int i = 0; i < 3; i++) { y[i] = e->TM[s][i * 3 + e->Z[par]] + b[s][i]; if (y[i] > mx) mx = y[i]; }
        for (int i = 0; i < 3; i++) p[i] = exp(y[i] - mx);
        int z = categorical(p, 3, slot_u(c->o->seed, 0, c2, c3, (uint32_t)t));
        if (e->Z[s] != z) changed[nch++] = s;
        e->Z[s] = z;
    }
    return nch;
}

/* BPP_C::getUpdateNode(changedZ) (bpp_c.cpp:620-642) */
static void update_node_changed(or_element *e, const int *changed, int nch, char *visited)
{
    for (int i = 0; i < e->N; i++) visited[i] = 1;
    for (int k = nch - 1; k >= 0; k--) {
        int j = e->par2[changed[k]];
        while (j != e->N) {
            if (!visited[j]) break;
            visited[j] = 0;
            for (int g = 0; g < e->GG; g++) { double *l = LAM(e, g, j); l[0] = l[1] = l[2] = l[3] = 0; }
            j = e->par2[j];
        }
    }
}

/* BPP_C::getUpdateNode(neut) (bpp_c.cpp:836-858) */
static void update_node_class(or_element *e, int neut, char *visited)
{
    for (int i = 0; i < e->N; i++) visited[i] = 1;
    for (int k = 0; k < e->n_nodes; k++) {
        int s = e->nodes[k];
        if (e->missing[s]) continue;
        if (!neut && (e->Z[s] == 0 || e->Z[s] == 2)) continue;
        if (neut && (e->Z[s] == 1 || e->Z[s] == 0)) continue;
        int j = e->par2[s];
        while (j != e->N) { if (!visited[j]) break; visited[j] = 0; j = e->par2[j]; }
    }
}

/* BPP_C::log_f_Xz(visited, lambda_tmp, neut, propos) (bpp_c.cpp:860-928) */
static double log_f_Xz_prop(chain_t *c, const char *visited, double *lt, int neut, double propos)
{
    or_element *e = c->e; int N = e->N;
    double (*Lprop)[16] = (double (*)[16])malloc(sizeof(double) * 16 * N);
    for (int k = 0; k < e->n_inodes; k++) {
        int s = e->inodes[k];
        if (visited[s]) continue;
        for (int cc = 0; cc < 2; cc++) {
            int chi = e->ch2[s][cc];
            if (e->missing[chi]) continue;
            const double *L;
            if (!neut) {
                if (e->Z[chi] == 1) { or_log_trans(c->m, e->dist2[chi] * propos, Lprop[chi]); L = Lprop[chi]; }
                else if (e->Z[chi] == 2) L = e->Lneut[chi]; else L = e->Lnull[chi];
            } else {
                if (e->Z[chi] == 2) { or_log_trans(c->m, e->dist2[chi] * propos, Lprop[chi]); L = Lprop[chi]; }
                else if (e->Z[chi] == 1) L = e->Lcons[chi]; else L = e->Lnull[chi];
            }
            for (int g = 0; g < e->GG; g++) {
                if (TG(e, g, chi) < 4) {
                    double o[4]; log_multi(L, lt + ((size_t)g * N + chi) * 4, o);
                    double *d = lt + ((size_t)g * N + s) * 4;
                    for (int j = 0; j < 4; j++) d[j] += o[j];
                }
            }
        }
    }
    double res = 0;
    for (int g = 0; g < e->GG; g++) {
        double *lr = lt + ((size_t)g * N + e->root) * 4;
        if (!visited[e->root]) for (int j = 0; j < 4; j++) lr[j] += c->logpi[j];
        res += log_exp_sum(lr, 4);
    }
    free(Lprop);
    return res;
}

/* BPP_C::sample_rate (bpp_c.cpp:930-1176), M = 1, adaptive, factor 0.5 */
static double sample_rate(chain_t *c, double old_rate, int neut, const char *visited, double *loglik_old)
{
    or_element *e = c->e; const or_options *o = c->o; const or_model *mod = c->m;
    int m = c->m_cur, N = e->N;
    double cur = old_rate, prior_a, prior_b;
    if (neut) { prior_a = o->nprior_a; prior_b = o->nprior_b; } else { prior_a = o->cprior_a; prior_b = o->cprior_b; }
    uint32_t c3 = tagword(c->ms, neut ? OR_TAG_R2 : OR_TAG_R1), c2 = (uint32_t)m;
    if (visited[e->root]) {
        if (c->resZ != 0) {
            stream_t sp = {o->seed, 2, c2, c3, 0};
            int capped = 0;
            cur = gamma_draw_c(&sp, prior_a, prior_b, &capped);
            if (capped) c->flags |= OR_FLAG_SAMPLER;
            if (neut) {
                if (o->ropt == 1 && cur < o->nlb) cur = old_rate;
                else if (o->ropt == 2 && cur < c->tc[m]) cur = old_rate;
            } else {
                if (o->ropt == 1 && cur > o->cub) cur = old_rate;
                else if (o->ropt == 2 && cur > c->tn[m + 1]) cur = old_rate;
            }
        } else cur = old_rate;
    } else {
        int skip = 0;
        double r = slot_u(o->seed, 0, c2, c3, 0), proposal;
        if (neut) {
            stream_t sg = {o->seed, 1, c2, c3, 0};
            int capped = 0;
            proposal = gamma_draw_c(&sg, cur / e->prop_n, e->prop_n, &capped);
            if (capped) c->flags |= OR_FLAG_SAMPLER;
            if (o->ropt == 1 && proposal < o->nlb) skip = 1;
            else if (o->ropt == 2 && proposal < c->tc[m]) skip = 1;
        } else {
            double ru = o->r1_shared_uniform ? r : slot_u(o->seed, 1, c2, c3, 0);
            double u = ru * (e->prop_c - 1 / e->prop_c) + 1 / e->prop_c;
            proposal = cur * u;
            if (o->ropt == 1 && proposal > o->cub) skip = 1;
            else if (o->ropt == 2 && proposal > c->tn[m + 1]) skip = 1;
        }
        if (!skip && (proposal < 1e-12 || proposal > 1e8)) {   /* the reference warns and skips (bpp_c.cpp:1024-1041) */
            skip = 1;
            c->flags |= OR_FLAG_PROPOSAL_RANGE;
        }
        if (!skip) {
            size_t GG = e->GG;
            double *lt = (double *)calloc(GG * N * 4, sizeof(double));
            for (int k = 0; k < e->n_nodes - 1; k++) {
                int s = e->nodes[k];
                if (!visited[s]) continue;
                for (size_t g = 0; g < GG; g++) memcpy(lt + (g * N + s) * 4, LAM(e, g, s), 4 * sizeof(double));
            }
            double loglik_new = log_f_Xz_prop(c, visited, lt, neut, proposal);
            double MH;
            if (neut)
                MH = loglik_new - *loglik_old + log_gamma_pdf(proposal, prior_a, prior_b) - log_gamma_pdf(cur, prior_a, prior_b)
#ifndef OR_MUTANT_NO_HASTINGS   /* planted mistake for the oracle's own tests: the Hastings terms dropped */
                     + log_gamma_pdf(cur, proposal / e->prop_n, e->prop_n) - log_gamma_pdf(proposal, cur / e->prop_n, e->prop_n)
#endif
                     ;
            else
                MH = loglik_new - *loglik_old + log_gamma_pdf(proposal, prior_a, prior_b) - log_gamma_pdf(cur, prior_a, prior_b)
                     + log(cur) - log(proposal);
            if (log(r) < MH) {
                cur = proposal; *loglik_old = loglik_new;
                for (int k = 0; k < e->n_inodes; k++) {
                    int s = e->inodes[k];
                    if (visited[s]) continue;
                    for (size_t g = 0; g < GG; g++) memcpy(LAM(e, g, s), lt + (g * N + s) * 4, 4 * sizeof(double));
                }
            }
            free(lt);
        }
    }
    if (old_rate != cur) {
        for (int k = 0; k < e->n_nodes - 1; k++) {
            int s = e->nodes[k];
            if (e->dist2[s] < 1e-10) continue;
            or_log_trans(mod, e->dist2[s] * cur, neut ? e->Lneut[s] : e->Lcons[s]);
        }
        if (neut) e->accept_n += 1; else e->accept_c += 1;
    }
    int af = o->adapt_freq;
    if ((m + 1) % af == 0) {
        double fac = 1 / sqrt((double)((m + 1) / af)), adj;
        if (fac > 0.5) fac = 0.5;
        if (neut) {
            double ar = (double)e->accept_n / af;
            adj = ar > 0.44 ? exp(fac) : (ar < 0.23 ? exp(-fac) : 1);
            e->prop_n = e->prop_n * adj; if (e->prop_n > 1) e->prop_n = 1;
            e->accept_n = 0;
        } else {
            double ar = (double)e->accept_c / af;
            adj = ar > 0.44 ? exp(fac) : (ar < 0.23 ? exp(-fac) : 1);
            e->prop_c = pow(e->prop_c, adj); if (e->prop_c < 1.0001) e->prop_c = 1.0001;
            e->accept_c = 0;
        }
    }
    return cur;
}

/* BPP_C::sample_transition (bpp_c.cpp:522-613) */
static void sample_transition(chain_t *c, double *gr, double *lr, double *lr2)
{
    or_element *e = c->e; const or_options *o = c->o;
    double nZ[3][3] = {{0}};
    for (int k = 0; k < e->n_nodes - 1; k++) { int s = e->nodes[k]; nZ[e->Z[s]][e->Z[e->par2[s]]] += 1; }
    uint32_t c3 = tagword(c->ms, OR_TAG_TRANS), c2 = (uint32_t)c->m_cur;
    int capped = 0;
#ifndef OR_MUTANT_SWAP_G_COUNTS   /* planted mistake for the oracle's own tests: g's counts swapped */
    *gr = beta_draw_c(o->seed, 0, 1, c2, c3, o->hyper_grate_a + nZ[1][0] + nZ[2][0], o->hyper_grate_b + nZ[0][0], &capped);
#else
    *gr = beta_draw_c(o->seed, 0, 1, c2, c3, o->hyper_grate_a + nZ[0][0], o->hyper_grate_b + nZ[1][0] + nZ[2][0], &capped);
#endif
    for (int k = 0; k < e->n_nodes - 1; k++) {
        int s = e->nodes[k];
        if (e->fixZ[s] == 1) nZ[e->Z[s]][e->Z[e->par2[s]]] -= 1;
    }
    *lr = beta_draw_c(o->seed, 2, 3, c2, c3, o->hyper_lrate_a + nZ[2][1], o->hyper_lrate_b + nZ[1][1], &capped);
    if (o->hyper_lrate2_a == 0) *lr2 = 0;
    else *lr2 = beta_draw_c(o->seed, 4, 5, c2, c3, o->hyper_lrate2_a + nZ[1][2], o->hyper_lrate2_b + nZ[2][2], &capped);
    if (capped) c->flags |= OR_FLAG_SAMPLER;
    for (int k = 0; k < e->n_nodes; k++) set_TM(e, e->nodes[k], *gr, *lr, *lr2, 0);
}

/* BPP_C::Gibbs (bpp_c.cpp:346-519) */
static void gibbs(chain_t *c)
{
    or_element *e = c->e;
    char visited[OR_MAXN];
    memset(visited, 0, sizeof visited);
    for (int g = 0; g < e->GG; g++) prune_site(e, g, visited, c->logpi);
    double log_lik_old = 0;
    int changed[OR_MAXN];
    for (int m = 0; m < c->T; m++) {
        c->m_cur = m;
        monitor(c, m, &log_lik_old);
        if (m == c->T - 1) break;
        for (int g = 0; g < e->GG; g++) sample_site(c, g);
        int nch = update_Z(c, changed);
        update_node_changed(e, changed, nch, visited);
        if (nch > 0) {
            log_lik_old = 0;
            for (int g = 0; g < e->GG; g++) { prune_site(e, g, visited, c->logpi); log_lik_old += log_exp_sum(LAM(e, g, e->root), 4); }
        }
        char vn[OR_MAXN], vc[OR_MAXN];
        update_node_class(e, 1, vn);
        update_node_class(e, 0, vc);
        c->tn[m + 1] = sample_rate(c, c->tn[m], 1, vn, &log_lik_old);
        c->tc[m + 1] = sample_rate(c, c->tc[m], 0, vc, &log_lik_old);
        sample_transition(c, &c->tg[m + 1], &c->tlr[m + 1], &c->tl2[m + 1]);
    }
}

/* BPP_C::Eval2 (bpp_c2.cpp:14-353) */
static double eval2(chain_t *c)
{
    or_element *e = c->e; const or_options *o = c->o;
    int N = e->N, T = c->T, nb = o->burnin, nm = o->mcmc;
    int *post = (int *)calloc(N, sizeof(int)), *unc = (int *)malloc(sizeof(int) * N), nu = 0;
    for (int s = 0; s < N; s++) e->Z[s] = c->Max_Z[s];
    for (int s = 0; s < N; s++) {
        for (int i = nb; i < T; i++) if (c->tZ[(size_t)i * N + s] == c->Max_Z[s]) post[s]++;
        if (post[s] < nm) unc[nu++] = s;
    }
    int cap = T + 1, nconf = 1;
    int *confrow = (int *)malloc(sizeof(int) * cap);   /* -1 means Max_Z */
    int *num = (int *)calloc(cap, sizeof(int));
    double *vc = (double *)calloc(cap, sizeof(double)), *vn = (double *)calloc(cap, sizeof(double));
    double *mc = (double *)calloc(cap, sizeof(double)), *mn = (double *)calloc(cap, sizeof(double));
    double *cov = (double *)calloc(cap, sizeof(double)), *sfl = (double *)calloc(cap, sizeof(double));
    confrow[0] = -1;
#define CONF(j, s) (confrow[j] < 0 ? c->Max_Z[s] : (int)c->tZ[(size_t)confrow[j] * N + (s)])
    for (int i = T - 1; i >= nb; i--) {
        int j = 0;
        for (; j < nconf; j++) {
            int k = 0;
            for (; k < nu; k++) if (c->tZ[(size_t)i * N + unc[k]] != CONF(j, unc[k])) break;
            if (k == nu) break;
        }
        if (j == nconf) { confrow[nconf] = i; nconf++; }
        num[j]++;
        vc[j] += pow(c->tc[i], 2); mc[j] += c->tc[i];
        if (c->resZ != 0) { vn[j] += pow(c->tn[i], 2); mn[j] += c->tn[i]; cov[j] += c->tc[i] * c->tn[i]; }
        sfl[j] += c->tf[i] / nm;
    }
    double VB = 0, logdet; int eff = 0;
    double prior_z[3] = {log(0.5), log(0.5), NEG_INF};
    for (int i = 0; i < nconf; i++) {
        double w = (double)num[i] / nm;
        if (w < 0.02) continue;
        double nZ[3][3] = {{0}};
        for (int k = 0; k < e->n_nodes - 1; k++) { int s = e->nodes[k]; nZ[CONF(i, s)][CONF(i, e->par2[s])] += 1; }
        double priorZ = lnbeta(o->hyper_grate_a + nZ[1][0] + nZ[2][0], o->hyper_grate_b + nZ[0][0]) - lnbeta(o->hyper_grate_a, o->hyper_grate_b);
        for (int k = 0; k < e->n_nodes - 1; k++) { int s = e->nodes[k]; if (e->fixZ[s] == 1) nZ[CONF(i, s)][CONF(i, e->par2[s])] -= 1; }
        priorZ += lnbeta(o->hyper_lrate_a + nZ[2][1], o->hyper_lrate_b + nZ[1][1]) - lnbeta(o->hyper_lrate_a, o->hyper_lrate_b);
        if (o->hyper_lrate2_a > 0) priorZ += lnbeta(o->hyper_lrate2_a + nZ[1][2], o->hyper_lrate2_b + nZ[2][2]) - lnbeta(o->hyper_lrate2_a, o->hyper_lrate2_b);
        priorZ += prior_z[CONF(i, e->root)];
        sfl[i] += num[i] * priorZ / nm;
        if (c->resZ == 0) logdet = log(vc[i] - pow(mc[i], 2) / num[i]) - log(num[i]);
        else {
            double a = vc[i] - pow(mc[i], 2) / num[i], b = vn[i] - pow(mn[i], 2) / num[i];
            double cv = cov[i] - mn[i] * mc[i] / num[i];
            logdet = log(a * b - cv * cv) - 2 * log(num[i]);
        }
        if (isnan(logdet)) continue;
        eff += num[i];
        VB -= w * log(num[i]);
        VB += sfl[i];
        VB += 0.5 * logdet * num[i] / nm;
    }
    if (eff == 0) {
        double varn = 0, varc = 0, cv = 0, meann = 0, meanc = 0;
        VB = 0;
        for (int i = 0; i < nconf; i++) {
            varc += vc[i]; meanc += mc[i];
            if (c->resZ != 0) { meann += mn[i]; varn += vn[i]; cv += cov[i]; }
        }
        double rn = meann / nm, rc = meanc / nm;
        for (int k = 0; k < e->n_nodes - 1; k++) {
            int s = e->nodes[k];
            or_log_trans(c->m, e->dist2[s] * rc, e->Lcons[s]);
            if (c->resZ != 0) or_log_trans(c->m, e->dist2[s] * rn, e->Lneut[s]);
        }
        for (int i = 0; i < nconf; i++) {
            double nZ[3][3] = {{0}};
            for (int k = 0; k < e->n_nodes - 1; k++) { int s = e->nodes[k]; nZ[CONF(i, s)][CONF(i, e->par2[s])] += 1; }
            double priorZ = lnbeta(o->hyper_grate_a + nZ[1][0] + nZ[2][0], o->hyper_grate_b + nZ[0][0]) - lnbeta(o->hyper_grate_a, o->hyper_grate_b);
            for (int k = 0; k < e->n_nodes - 1; k++) { int s = e->nodes[k]; if (e->fixZ[s] == 1) nZ[CONF(i, s)][CONF(i, e->par2[s])] -= 1; }
            priorZ += lnbeta(o->hyper_lrate_a + nZ[2][1], o->hyper_lrate_b + nZ[1][1]) - lnbeta(o->hyper_lrate_a, o->hyper_lrate_b);
            if (o->hyper_lrate2_a > 0) priorZ += lnbeta(o->hyper_lrate2_a + nZ[1][2], o->hyper_lrate2_b + nZ[2][2]) - lnbeta(o->hyper_lrate2_a, o->hyper_lrate2_b);
            priorZ += prior_z[CONF(i, e->root)];
            for (int s = 0; s < N; s++) e->Z[s] = CONF(i, s);
            zero_internal(e);
            char visited[OR_MAXN]; memset(visited, 0, sizeof visited);
            sfl[i] = 0;
            for (int g = 0; g < e->GG; g++) { prune_site(e, g, visited, c->logpi); sfl[i] += log_exp_sum(LAM(e, g, e->root), 4); }
            sfl[i] += priorZ;
        }
        VB = log_exp_sum(sfl, nconf);
        VB += log_gamma_pdf(rc, o->cprior_a, o->cprior_b);
        if (c->resZ != 0) VB += log_gamma_pdf(rn, o->nprior_a, o->nprior_b);
        if (c->resZ == 0) {
            varc -= pow(meanc, 2) / nm;
            logdet = log(varc) - log(nm);
            VB += 0.5 * logdet + 0.5 * log(2 * M_PI);
        } else {
            varn -= pow(meann, 2) / nm; varc -= pow(meanc, 2) / nm; cv -= meann * meanc / nm;
            logdet = log(varc * varn - cv * cv) - 2 * log(nm);
            VB += 0.5 * logdet + log(2 * M_PI);
        }
    } else {
        if (c->resZ != 0) VB = VB * ((double)nm / eff) + log(nm) + log(2 * M_PI) + 1;
        else VB = VB * ((double)nm / eff) + log(nm) + 0.5 * (log(2 * M_PI) + 1);
    }
#undef CONF
    free(post); free(unc); free(confrow); free(num); free(vc); free(vn); free(mc); free(mn); free(cov); free(sfl);
    return VB;
}

static int cmp_d(const void *a, const void *b) { double x = *(const double *)a, y = *(const double *)b; return x < y ? -1 : (x > y); }

/* BPP_C::Output_init (bpp_c2.cpp:739-785): medians over the whole trace and
 * post-burn-in Z frequencies (missing nodes: column 0 set to MCMC) */
static void output_init(chain_t *c, or_chain_out *out)
{
    or_element *e = c->e; int T = c->T, N = e->N;
    double *tmp = (double *)malloc(sizeof(double) * T);
    double *src[5] = {c->tn, c->tc, c->tg, c->tlr, c->tl2};
    for (int k = 0; k < 5; k++) { memcpy(tmp, src[k], sizeof(double) * T); qsort(tmp, T, sizeof(double), cmp_d); out->med[k] = tmp[T / 2]; }
    free(tmp);
    if (!out->postZ) return;
    for (int s = 0; s < N; s++) {
        int cnt[4] = {0};
        for (int i = c->o->burnin; i < T; i++) cnt[c->tZ[(size_t)i * N + s] + 1]++;
        if (e->missing[s]) cnt[0] = c->o->mcmc;
        for (int k = 0; k < 4; k++) out->postZ[s * 4 + k] = (double)cnt[k] / c->o->mcmc;
    }
}

int or_run_element(const or_model *m, const or_options *o, or_element *e, or_chain_out *outs)
{
    if (e->filtered) return 1;
    int T = o->burnin + o->mcmc, N = e->N;
    static const int order[3] = {0, 2, 1};
    e->accept_n = e->accept_c = 0;
    for (int ms = 0; ms < 3; ms++) {
        chain_t c; memset(&c, 0, sizeof c);
        c.m = m; c.o = o; c.e = e; c.resZ = order[ms]; c.ms = ms; c.T = T;
        for (int j = 0; j < 4; j++) c.logpi[j] = log(m->pi[j]);
        or_chain_out *out = &outs[ms];
        double *own[7] = {0}; signed char *ownZ = NULL;
        double **dst[7] = {&c.tl, &c.tf, &c.tn, &c.tc, &c.tg, &c.tlr, &c.tl2};
        double *given[7] = {out->loglik, out->full, out->nrate, out->crate, out->grate, out->lrate, out->l2rate};
        for (int k = 0; k < 7; k++) {
            if (given[k]) *dst[k] = given[k];
            else { own[k] = (double *)calloc(T, sizeof(double)); *dst[k] = own[k]; }
        }
        if (out->Z) c.tZ = out->Z; else { ownZ = (signed char *)calloc((size_t)T * N, 1); c.tZ = ownZ; }
        init_mcmc(&c);
        gibbs(&c);
        out->max_loglik = c.MaxLoglik; out->max_m = c.Max_m;
        out->flags = c.flags;
        for (int s = 0; s < N; s++) out->max_Z[s] = c.Max_Z[s];
        out->vb = eval2(&c);
        out->failure = isnan(out->vb) ? 1 : 0;
        output_init(&c, out);
        out->accept_n = e->accept_n; out->accept_c = e->accept_c;
        for (int k = 0; k < 7; k++) free(own[k]);
        free(ownZ);
    }
    return 0;
}

long or_run_batch(const or_model *m, const or_options *o, or_element **es, int n, int nthreads, double *max_loglik)
{
    long sweeps = 0;
    int T = o->burnin + o->mcmc;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(guided) num_threads(nthreads) reduction(+ : sweeps)
#endif
    for (int i = 0; i < n; i++) {
        if (es[i]->filtered) continue;
        or_chain_out outs[3]; memset(outs, 0, sizeof outs);
        or_run_element(m, o, es[i], outs);
        for (int k = 0; k < 3; k++) if (max_loglik) max_loglik[i * 3 + k] = outs[k].max_loglik;
        sweeps += 3L * (T - 1);
    }
    return sweeps;
}

/* Test hook: one Z update on a bare tree with given per-branch
 * log-emissions E[N*3] and transition probabilities g, l, l2.
 * nodes: post-order (root last); parent: [N] with parent[root] = N. */
int or_test_update_Z(int N, const int *nodes, int n_nodes, const int *parent, const int *fixZ, const double *E,
                     double g, double l, double l2, uint64_t seed, int m, int ms, int *Zout)
{
    or_element *e = (or_element *)calloc(1, sizeof(or_element));
    or_options o; memset(&o, 0, sizeof o); o.seed = seed;
    e->N = N; e->n_nodes = n_nodes; e->root = nodes[n_nodes - 1];
    for (int k = 0; k < n_nodes; k++) e->nodes[k] = nodes[k];
    for (int s = 0; s < N; s++) {
        e->par2[s] = parent[s]; e->fixZ[s] = fixZ[s]; e->Z[s] = 1;
        for (int j = 0; j < 3; j++) e->log_emission[s][j] = E[s * 3 + j];
        for (int j = 0; j < 9; j++) e->TM[s][j] = 0;
    }
    for (int k = 0; k < n_nodes; k++) set_TM(e, nodes[k], g, l, l2, 1);
    chain_t c; memset(&c, 0, sizeof c);
    c.e = e; c.o = &o; c.ms = ms; c.m_cur = m;
    int changed[OR_MAXN];
    update_Z_core(&c, changed);
    for (int s = 0; s < N; s++) Zout[s] = e->Z[s];
    free(e);
    return 0;
}

/* Test hook: Eval2 (bpp_c2.cpp:14-353) over caller-supplied traces of model
 * slot ms (0 = M0, 1 = M1, 2 = M2) of element e: full, nrate, crate [T] and
 * Z [T*N] with T = burnin + mcmc, and the configuration max_Z [N].  The
 * element's fixZ and caches are set up as initMCMC does for that model. */
double or_test_eval2(const or_model *m, const or_options *o, or_element *e, int ms, const double *full,
                     const double *nrate, const double *crate, const signed char *Z, const int *max_Z)
{
    if (e->filtered) return NAN;
    static const int order[3] = {0, 2, 1};
    int T = o->burnin + o->mcmc, N = e->N;
    chain_t c; memset(&c, 0, sizeof c);
    c.m = m; c.o = o; c.e = e; c.resZ = order[ms]; c.ms = ms; c.T = T;
    for (int j = 0; j < 4; j++) c.logpi[j] = log(m->pi[j]);
    double *buf = (double *)calloc((size_t)7 * T, sizeof(double));
    c.tl = buf; c.tf = buf + T; c.tn = buf + 2 * T; c.tc = buf + 3 * T;
    c.tg = buf + 4 * T; c.tlr = buf + 5 * T; c.tl2 = buf + 6 * T;
    c.tZ = (signed char *)malloc((size_t)T * N);
    init_mcmc(&c);
    memcpy(c.tf, full, sizeof(double) * T);
    memcpy(c.tn, nrate, sizeof(double) * T);
    memcpy(c.tc, crate, sizeof(double) * T);
    memcpy(c.tZ, Z, (size_t)T * N);
    for (int s = 0; s < N; s++) c.Max_Z[s] = max_Z[s];
    double vb = eval2(&c);
    free(buf); free(c.tZ);
    return vb;
}

/* Test hook: the ancestral draw of one site (BPP_C::Update_Tg(tosample=true),
 * bpp_c.cpp:672-733) at fixed Z and rates (r1 conserved, r2 accelerated),
 * with the uniforms of sweep m_sweep of model slot ms.  states: [N], the
 * sampled base of every internal node (leaves keep their encoded Tg). */
int or_test_sample_site(const or_model *m, or_element *e, const int *Z, double r1, double r2, uint64_t seed,
                        int m_sweep, int ms, int site, int *states)
{
    if (e->filtered || site < 0 || site >= e->GG) return 1;
    or_options o; memset(&o, 0, sizeof o); o.seed = seed;
    model_caches(m, e, r2, r1);
    for (int s = 0; s < e->N; s++) e->Z[s] = Z[s];
    zero_internal(e);
    chain_t c; memset(&c, 0, sizeof c);
    c.m = m; c.o = &o; c.e = e; c.ms = ms; c.m_cur = m_sweep;
    for (int j = 0; j < 4; j++) c.logpi[j] = log(m->pi[j]);
    char visited[OR_MAXN] = {0};
    prune_site(e, site, visited, c.logpi);
    sample_site(&c, site);
    for (int s = 0; s < e->N; s++) states[s] = TG(e, site, s);
    return 0;
}

/* Gene-tree pruning at fixed species-branch Z and rates: GTree::Update_Tg
 * (tosample = false, src/PhyloAcc-GT/genetree.cpp:1330-1426).  Every gene
 * branch's message is composed piece by piece, bottom-up, one log_multi per
 * piece it spends in a species branch, at that branch's rate; pieces shorter
 * than 1e-8 are the identity (getlogTMc, src/PhyloAcc-GT/bpp.cpp:221-235).
 * Gene leaves are the species 0..S-1; internal gene ids are post-order. */
double or_gt_loglik(const or_model *m, int S, const int *gchildren, int K, const int *seg_gene, const int *seg_sp,
                    const double *seg_len, const int *Z, double r1, double r2, const char *const *rows, long start,
                    int GG, int gapchar, double gap_prop)
{
    int N = 2 * S - 1;
    double logpi[4];
    for (int j = 0; j < 4; j++) logpi[j] = log(m->pi[j]);
    /* per piece transition (log P transposed) */
    double (*L)[16] = (double (*)[16])malloc(sizeof(double) * 16 * (K > 0 ? K : 1));
    for (int k = 0; k < K; k++) {
        int z = Z[seg_sp[k]];
        double rate = z == 2 ? r2 : (z == 1 ? r1 : 1.0);
        if (seg_len[k] < 1e-8) log_identity(L[k]);
        else or_log_trans(m, seg_len[k] * rate, L[k]);
    }
    double (*lam)[4] = (double (*)[4])malloc(sizeof(double) * 4 * N);
    int *gap = (int *)malloc(sizeof(int) * N);
    /* missing genes: a leaf with more than gap_prop*GG gap characters, an
     * internal gene node whose children are both missing (bpp_c.cpp:34-60);
     * the pruning skips them (bpp_c.cpp:1130, 1156, 1186) */
    int *miss = (int *)calloc(N, sizeof(int));
    for (int s = 0; s < S; s++) {
        int ng = 0;
        for (int g = 0; g < GG; g++) ng += tolower((unsigned char)rows[s][start + g]) == gapchar;
        miss[s] = ng > gap_prop * GG;
    }
    for (int s = S; s < N; s++) miss[s] = miss[gchildren[2 * s]] && miss[gchildren[2 * s + 1]];
    double total = 0;
    for (int g = 0; g < GG; g++) {
        for (int s = 0; s < S; s++) {
            int Tg; unsigned char code;
            encode_char(tolower((unsigned char)rows[s][start + g]), gapchar, lam[s], &Tg, &code);
            gap[s] = Tg >= 4;
        }
        for (int s = S; s < N; s++) {
            for (int j = 0; j < 4; j++) lam[s][j] = 0;
            int a = gchildren[2 * s], b = gchildren[2 * s + 1];
            gap[s] = gap[a] && gap[b];
            for (int q = 0; q < 2; q++) {
                int x = q ? b : a;
                if (gap[x] || miss[x]) continue;
                double msg[4], tmp[4];
                memcpy(msg, lam[x], sizeof msg);
                for (int k = 0; k < K; k++) {
                    if (seg_gene[k] != x) continue;
                    log_multi(L[k], msg, tmp);
                    memcpy(msg, tmp, sizeof msg);
                }
                for (int j = 0; j < 4; j++) lam[s][j] += msg[j];
            }
        }
        double y[4];
        for (int j = 0; j < 4; j++) y[j] = lam[N - 1][j] + logpi[j];
        total += log_exp_sum(y, 4);
    }
    free(L); free(lam); free(gap); free(miss);
    return total;
}

/* ------------------------------------------- GT gene tree in the species tree -- */

/* Per-site gene-tree log-likelihoods (the summands of or_gt_loglik) for a
 * gene tree in any numbering: gchildren [N*2] (-1 at leaves), root groot.
 * Missing genes and gaps as or_gt_loglik, over the element's GG sites. */
void or_gt_site_logliks(const or_model *m, int S, const int *gchildren, int groot, int K, const int *seg_gene,
                        const int *seg_sp, const double *seg_len, const int *Z, double r1, double r2,
                        const char *const *rows, long start, int GG, int gapchar, double gap_prop, double *site)
{
    int N = 2 * S - 1;
    double logpi[4];
    for (int j = 0; j < 4; j++) logpi[j] = log(m->pi[j]);
    double (*L)[16] = (double (*)[16])malloc(sizeof(double) * 16 * (K > 0 ? K : 1));
    for (int k = 0; k < K; k++) {
        int z = Z[seg_sp[k]];
        double rate = z == 2 ? r2 : (z == 1 ? r1 : 1.0);
        if (seg_len[k] < 1e-8) log_identity(L[k]);
        else or_log_trans(m, seg_len[k] * rate, L[k]);
    }
    /* post-order of the gene tree from its root */
    int order[OR_MAXN], no = 0, stk[OR_MAXN], vis[OR_MAXN], ns = 0;
    stk[ns] = groot; vis[ns++] = 0;
    while (ns > 0) {
        int v = stk[ns - 1];
        if (v >= S && vis[ns - 1] == 0) {
            vis[ns - 1] = 1;
            stk[ns] = gchildren[2 * v + 1]; vis[ns++] = 0;
            stk[ns] = gchildren[2 * v]; vis[ns++] = 0;
        } else {
            order[no++] = v;
            ns--;
        }
    }
    double (*lam)[4] = (double (*)[4])malloc(sizeof(double) * 4 * N);
    int *gap = (int *)malloc(sizeof(int) * N), *miss = (int *)calloc(N, sizeof(int));
    for (int s = 0; s < S; s++) {
        int ng = 0;
        for (int g = 0; g < GG; g++) ng += tolower((unsigned char)rows[s][start + g]) == gapchar;
        miss[s] = ng > gap_prop * GG;
    }
    for (int q = 0; q < no; q++) {
        int s = order[q];
        if (s >= S) miss[s] = miss[gchildren[2 * s]] && miss[gchildren[2 * s + 1]];
    }
    for (int g = 0; g < GG; g++) {
        for (int s = 0; s < S; s++) {
            int Tg; unsigned char code;
            encode_char(tolower((unsigned char)rows[s][start + g]), gapchar, lam[s], &Tg, &code);
            gap[s] = Tg >= 4;
        }
        for (int q = 0; q < no; q++) {
            int s = order[q];
            if (s < S) continue;
            for (int j = 0; j < 4; j++) lam[s][j] = 0;
            int a = gchildren[2 * s], b = gchildren[2 * s + 1];
            gap[s] = gap[a] && gap[b];
            for (int c = 0; c < 2; c++) {
                int x = c ? b : a;
                if (gap[x] || miss[x]) continue;
                double msg[4], tmp[4];
                memcpy(msg, lam[x], sizeof msg);
                for (int k = 0; k < K; k++) {
                    if (seg_gene[k] != x) continue;
                    log_multi(L[k], msg, tmp);
                    memcpy(msg, tmp, sizeof msg);
                }
                for (int j = 0; j < 4; j++) lam[s][j] += msg[j];
            }
        }
        double y[4];
        for (int j = 0; j < 4; j++) y[j] = lam[groot][j] + logpi[j];
        site[g] = log_exp_sum(y, 4);
    }
    free(L); free(lam); free(gap); free(miss);
}

/* GTree::findSp (src/PhyloAcc-GT/genetree.cpp:2521-2534): from gene node v
 * down its first children to a leaf, whose species is the leaf id; then up
 * the species tree while the parent species is lower than v. */
int or_gt_species(const or_sptree *sp, const int *gch, const double *gh, int v)
{
    int s = v;
    while (gch[2 * s] != -1) s = gch[2 * s];
    int p = sp->parent[s];
    while (p < sp->N && sp->height[p] < gh[v]) { s = p; p = sp->parent[s]; }
    return s;
}

/* Is gene lineage v (from v up to its parent gene node, or up through the
 * root species for the gene root) in species branch s?  The lineage runs
 * through the species from findSp(v) up to findSp(parent), both included
 * (GTree::parent_gene2[s] holds exactly these lineages as keys). */
static int gt_present(const or_sptree *sp, const int *gch, const int *gpar, const double *gh, int v, int s)
{
    int a = or_gt_species(sp, gch, gh, v);
    int top = gpar[v] < sp->N ? or_gt_species(sp, gch, gh, gpar[v]) : sp->N - 1;
    for (;;) {
        if (a == s) return 1;
        if (a == top || a >= sp->N - 1) return 0;
        a = sp->parent[a];
    }
}

/* GTree::priorTree (genetree.cpp:560-584): for every internal species s, nn
 * lineages entering from below coalesce at the gene nodes born in s in
 * height order, each at rate nn(nn-1)/theta_s with density 2/theta_s; the
 * lineages left at the top survive to the parent species. */
double or_gt_prior(const or_sptree *sp, const int *gch, const int *gpar, const double *gh, int groot)
{
    int S = sp->S, N = sp->N;
    double logP = 0;
    for (int s = S; s < N; s++) {
        int nn = 0, nb = 0;
        double born[OR_MAXN];
        for (int v = 0; v < N; v++) {
            if (v >= S && gch[2 * v] == -1) continue;
            if (v != groot && gpar[v] < 0) continue;
            if (!gt_present(sp, gch, gpar, gh, v, s)) continue;
            if (v >= S && or_gt_species(sp, gch, gh, v) == s) born[nb++] = gh[v];
            else nn++;
        }
        if (nn == 0) continue;
        for (int i = 1; i < nb; i++)           /* insertion sort: ascending heights */
            for (int j = i; j > 0 && born[j - 1] > born[j]; j--) { double t = born[j]; born[j] = born[j - 1]; born[j - 1] = t; }
        double cur = sp->height[s];
        for (int i = 0; i < nb; i++) {
            logP -= (born[i] - cur) * (nn - 1) * nn / sp->theta[s];
            logP += log(2.0) - log(sp->theta[s]);
            nn--;
            cur = born[i];
        }
        if (nn > 1 && sp->parent[s] != N)
            logP -= (sp->height[sp->parent[s]] - cur) * (nn - 1) * nn / sp->theta[s];
    }
    return logP;
}

/* GTree::get_logpZ1 (genetree.cpp:1156-1190): every tip, then every gene node
 * born in an internal species s, scores log P(missing | Z_s) with
 * P(missing | Z = 1) = consToMis and P(missing | Z != 1) = nconsToMis. */
double or_gt_logpz(const or_sptree *sp, const int *gch, const double *gh, const int *gmiss, const int *Z,
                   double c2m, double n2m)
{
    int S = sp->S, N = sp->N;
    double lp = 0;
    for (int s = 0; s < S; s++)
        lp += gmiss[s] ? (Z[s] == 1 ? log(c2m) : log(n2m)) : (Z[s] == 1 ? log(1 - c2m) : log(1 - n2m));
    for (int v = S; v < N; v++) {
        if (gch[2 * v] == -1) continue;
        int s = or_gt_species(sp, gch, gh, v);
        if (s < S || gh[v] < sp->height[s]) continue;
        lp += gmiss[v] ? (Z[s] == 1 ? log(c2m) : log(n2m)) : (Z[s] == 1 ? log(1 - c2m) : log(1 - n2m));
    }
    return lp;
}

/* The pieces every gene branch spends in species branches: lineage v from
 * max(h(v), bottom of s) to min(h(parent), top of s) in each species s on its
 * path (GTree::Update_Tg / Update_Lambda compose one transition per piece).
 * Returns K; the arrays hold up to N * (depth + 1) pieces. */
int or_gt_pieces(const or_sptree *sp, const int *gch, const int *gpar, const double *gh, int groot, int *seg_gene,
                 int *seg_sp, double *seg_len)
{
    int S = sp->S, N = sp->N, K = 0;
    for (int v = 0; v < N; v++) {
        if (v == groot || gpar[v] < 0 || gpar[v] >= N) continue;
        if (v >= S && gch[2 * v] == -1) continue;
        int s = or_gt_species(sp, gch, gh, v), top = or_gt_species(sp, gch, gh, gpar[v]);
        double lo = gh[v];
        for (;;) {
            double bottom = sp->height[s] > lo ? sp->height[s] : lo;
            double hi = (s == top || sp->parent[s] >= N) ? gh[gpar[v]] : sp->height[sp->parent[s]];
            seg_gene[K] = v; seg_sp[K] = s; seg_len[K] = hi - bottom; K++;
            if (s == top || sp->parent[s] >= N) break;
            s = sp->parent[s];
        }
    }
    return K;
}

/* GTree::Sample_BranchLen (genetree.cpp:1658-1965) on a gene tree held as
 * topology + heights: gene node gnode's height moves uniformly in a window of
 * half-width delta * theta_ss / 2 (ss: gnode's species) clipped to
 * [max(bottom of css, h(cnode)), upper1], where cnode is the taller child, css
 * the lowest species on cnode's path that the sibling's lineage also reaches,
 * and upper1 the parent gene node when that coalesces in ss, else the top of
 * ss (the gene root: 10 theta_root / 2, or h(root) + 5 theta_root / 2 when
 * that is lower).  Metropolis-Hastings on the prefix log-likelihood of the
 * WL indicator, the missing-gene term (when logpz[indicator] != 0), the
 * multispecies-coalescent prior and the clipped window's proposal ratio.
 * Draws: tag GT_BRANCH, sub-stream slot, sweep m: the height's uniform
 * (floored at 1e-5) at 0, the decision's at 1; none when the window is empty.
 * gh, loglik [2], logpz [2] and *prior are updated on acceptance.  Returns 1
 * accepted, 0 rejected or empty window, -1 when no species holds both
 * children's lineages (the reference throws). */
int or_gt_branch_len(const or_model *m, double piA, const or_sptree *sp, const int *gch, const int *gpar, double *gh,
                     const int *gmiss, int groot, const int *Z, double c_rate, double n_rate, double c2m, double n2m,
                     int gnode, double delta, int indicator, const int *lens, double *loglik, double *logpz,
                     double *prior, uint64_t seed, int m_sweep, int slot, const char *const *rows, long start, int GG,
                     int gapchar, double gap_prop)
{
    int N = sp->N;
    const double *H = sp->height;
    int c0 = gch[2 * gnode], c1 = gch[2 * gnode + 1], cnode, sib;
    if (gh[gnode] - gh[c0] < gh[gnode] - gh[c1]) { cnode = c0; sib = c1; }
    else { cnode = c1; sib = c0; }
    int ss1 = or_gt_species(sp, gch, gh, cnode), ss = ss1, css = ss1, spp = ss, found = 0;
    while (spp < N && H[spp] < gh[gnode]) {
        ss = spp;
        spp = sp->parent[ss];
        if (!found && gt_present(sp, gch, gpar, gh, sib, ss)) { css = ss; found = 1; }
    }
    if (!found) return -1;
    int pg = gpar[gnode];
    double upper1;
    if (pg < N && or_gt_species(sp, gch, gh, pg) == ss) upper1 = gh[pg];
    else upper1 = spp < N ? H[spp] : INFINITY;
    if (gnode == groot) {
        upper1 = 10 * sp->theta[N - 1] / 2;
        if (upper1 < gh[groot]) upper1 = gh[groot] + 5 * sp->theta[N - 1] / 2;
    }
    double lower1 = H[css] > gh[cnode] ? H[css] : gh[cnode];
    double half = delta * sp->theta[ss] / 2;
    double upper = fmin(upper1, gh[gnode] + half), lower = fmax(lower1, gh[gnode] - half);
    double difL = upper - lower;
    if (difL == 0) return 0;
    uint32_t c3 = tagword(0, OR_TAG_GT_BRANCH), c2 = (uint32_t)m_sweep;
    double u = slot_u(seed, (uint32_t)slot, c2, c3, 0);
    double prop = fmax(u, 1e-5) * difL + lower;
    double r = slot_u(seed, (uint32_t)slot, c2, c3, 1);
    double old = gh[gnode];
    gh[gnode] = prop;
    /* the proposal's prefix log-likelihoods (GG sites, Wang-Landau lens) */
    int cap = N * (N + 1);
    int *sg = (int *)malloc(sizeof(int) * cap), *ssp = (int *)malloc(sizeof(int) * cap);
    double *sl = (double *)malloc(sizeof(double) * cap), *site = (double *)malloc(sizeof(double) * (GG > 0 ? GG : 1));
    int K = or_gt_pieces(sp, gch, gpar, gh, groot, sg, ssp, sl);
    or_model mm;
    if (piA > 0) gt_model_at(m, piA, &mm); else mm = *m;
    or_gt_site_logliks(&mm, sp->S, gch, groot, K, sg, ssp, sl, Z, c_rate, n_rate, rows, start, GG, gapchar, gap_prop,
                       site);
    double lnew[2] = {0, 0};
    for (int g = 0; g < lens[indicator]; g++) {
        lnew[0] += site[g];
        if (g < lens[1]) lnew[1] += site[g];
    }
    free(sg); free(ssp); free(sl); free(site);
    double lpz_new = or_gt_logpz(sp, gch, gh, gmiss, Z, c2m, n2m);
    double cur_den = -log(upper - lower);
    double lower2 = fmax(lower1, prop - half), upper2 = fmin(upper1, prop + half);
    double pro_den = -log(upper2 - lower2);
    double d = lnew[indicator] - loglik[indicator] + pro_den - cur_den;
    if (logpz[indicator] != 0) d += lpz_new - logpz[indicator];
    double prior_new = or_gt_prior(sp, gch, gpar, gh, groot);
    d += prior_new - *prior;
    if (log(r) < d) {
        loglik[0] = lnew[0];
        loglik[1] = lnew[1];
        if (logpz[1] == 0) logpz[0] = lpz_new;
        else logpz[0] = logpz[1] = lpz_new;
        *prior = prior_new;
        return 1;
    }
    gh[gnode] = old;
    return 0;
}

/* Shared set-up of one gene tree for per-site work: the post-order from the
 * root, each piece's log transition (species rate applied) and the missing
 * genes (more than gap_prop of the GG sites are gaps; an internal gene when
 * both children are). */
typedef struct {
    int S, N, groot, no, K;
    int order[OR_MAXN], miss[OR_MAXN];
    int *sg, *ssp;
    double *sl, (*L)[16];
    double logpi[4];
} gt_ctx;

static void gt_ctx_init(gt_ctx *c, const or_model *m, const or_sptree *sp, const int *gch, const int *gpar,
                        const double *gh, int groot, const int *Z, double c_rate, double n_rate,
                        const char *const *rows, long start, int GG, int gapchar, double gap_prop)
{
    int N = sp->N, S = sp->S, cap = N * (N + 1);
    c->S = S; c->N = N; c->groot = groot;
    c->sg = (int *)malloc(sizeof(int) * cap);
    c->ssp = (int *)malloc(sizeof(int) * cap);
    c->sl = (double *)malloc(sizeof(double) * cap);
    c->K = or_gt_pieces(sp, gch, gpar, gh, groot, c->sg, c->ssp, c->sl);
    c->L = (double (*)[16])malloc(sizeof(double) * 16 * (c->K > 0 ? c->K : 1));
    for (int k = 0; k < c->K; k++) {
        int z = Z[c->ssp[k]];
        double rate = z == 2 ? n_rate : (z == 1 ? c_rate : 1.0);
        if (c->sl[k] < 1e-8) log_identity(c->L[k]);
        else or_log_trans(m, c->sl[k] * rate, c->L[k]);
    }
    for (int j = 0; j < 4; j++) c->logpi[j] = log(m->pi[j]);
    int stk[OR_MAXN], vis[OR_MAXN], ns = 0;
    c->no = 0;
    stk[ns] = groot; vis[ns++] = 0;
    while (ns > 0) {
        int v = stk[ns - 1];
        if (v >= S && vis[ns - 1] == 0) {
            vis[ns - 1] = 1;
            stk[ns] = gch[2 * v + 1]; vis[ns++] = 0;
            stk[ns] = gch[2 * v]; vis[ns++] = 0;
        } else {
            c->order[c->no++] = v;
            ns--;
        }
    }
    for (int s = 0; s < N; s++) c->miss[s] = 0;
    for (int s = 0; s < S; s++) {
        int ng = 0;
        for (int g = 0; g < GG; g++) ng += tolower((unsigned char)rows[s][start + g]) == gapchar;
        c->miss[s] = ng > gap_prop * GG;
    }
    for (int q = 0; q < c->no; q++) {
        int s = c->order[q];
        if (s >= S) c->miss[s] = c->miss[gch[2 * s]] && c->miss[gch[2 * s + 1]];
    }
}

static void gt_ctx_free(gt_ctx *c) { free(c->sg); free(c->ssp); free(c->sl); free(c->L); }

/* Site g's log partials lam[v] of every node on the tree and gap[v] (the
 * subtree holds no base), messages composed piece by piece (Update_Tg with
 * tosample = false); tg[v] of the leaves: 0-3 a/c/g/t, -1 ambiguous, >= 4 no
 * base (the GT encoding, src/PhyloAcc-GT/bpp_c.hpp:340-384). */
static void gt_site_partials(const gt_ctx *c, const int *gch, const char *const *rows, long start, int g, int gapchar,
                             double (*lam)[4], int *gap, int *tg)
{
    for (int s = 0; s < c->S; s++) {
        unsigned char code;
        encode_char(tolower((unsigned char)rows[s][start + g]), gapchar, lam[s], &tg[s], &code);
        gap[s] = tg[s] >= 4;
    }
    for (int q = 0; q < c->no; q++) {
        int s = c->order[q];
        if (s < c->S) continue;
        for (int j = 0; j < 4; j++) lam[s][j] = 0;
        int a = gch[2 * s], b = gch[2 * s + 1];
        gap[s] = gap[a] && gap[b];
        for (int k2 = 0; k2 < 2; k2++) {
            int x = k2 ? b : a;
            if (gap[x] || c->miss[x]) continue;
            double msg[4], tmp[4];
            memcpy(msg, lam[x], sizeof msg);
            for (int k = 0; k < c->K; k++) {
                if (c->sg[k] != x) continue;
                log_multi(c->L[k], msg, tmp);
                memcpy(msg, tmp, sizeof msg);
            }
            for (int j = 0; j < 4; j++) lam[s][j] += msg[j];
        }
    }
}

/* GTree::Update_Tg with tosample = true (src/PhyloAcc-GT/genetree.cpp:
 * 1416-1532): for sites g < len, the gene root's base from pi * exp(lam_root),
 * then down the tree each gene node's base from P(parent base -> .) *
 * exp(lam) over its whole branch (the reference steps through the species
 * boundaries; the composed draw has the same law).  Gene nodes that are
 * missing genes or hold no base at the site get 4; leaves keep their
 * encoding (0-3, -1 ambiguous, 4 no base).  Tg [N * len], node-major.
 * Draws: tag GT_HIST, sub-stream g, sweep m, slot = node id. */
void or_gt_sample_history(const or_model *m, double piA, const or_sptree *sp, const int *gch, const int *gpar,
                          const double *gh, int groot, const int *Z, double c_rate, double n_rate, uint64_t seed,
                          int m_sweep, int len, const char *const *rows, long start, int GG, int gapchar,
                          double gap_prop, signed char *Tg)
{
    or_model mm;
    if (piA > 0) gt_model_at(m, piA, &mm); else mm = *m;
    gt_ctx c;
    gt_ctx_init(&c, &mm, sp, gch, gpar, gh, groot, Z, c_rate, n_rate, rows, start, GG, gapchar, gap_prop);
    int N = sp->N;
    double (*lam)[4] = (double (*)[4])malloc(sizeof(double) * 4 * N);
    double (*P)[16] = (double (*)[16])malloc(sizeof(double) * 16 * N);
    int gap[OR_MAXN], tg[OR_MAXN];
    /* each branch's transition over its whole substitution length */
    for (int v = 0; v < N; v++) {
        double len_v = 0;
        for (int k = 0; k < c.K; k++)
            if (c.sg[k] == v && c.sl[k] >= 1e-8) {
                int z = Z[c.ssp[k]];
                len_v += c.sl[k] * (z == 2 ? n_rate : (z == 1 ? c_rate : 1.0));
            }
        double LP[16];
        if (len_v < 1e-8) log_identity(LP); else or_log_trans(&mm, len_v, LP);
        for (int q = 0; q < 16; q++) P[v][q] = exp(LP[q]);
    }
    uint32_t c3 = tagword(0, OR_TAG_GT_HIST), c2 = (uint32_t)m_sweep;
    for (int g = 0; g < len; g++) {
        gt_site_partials(&c, gch, rows, start, g, gapchar, lam, gap, tg);
        for (int v = 0; v < N; v++) Tg[(size_t)v * len + g] = 4;
        for (int s = 0; s < c.S; s++) Tg[(size_t)s * len + g] = (signed char)(tg[s] >= 4 ? 4 : tg[s]);
        /* pre-order: the reverse of the post-order */
        for (int q = c.no - 1; q >= 0; q--) {
            int v = c.order[q];
            if (v < c.S) continue;
            double p[4], mx = NEG_INF;
            if (v == groot) {
                for (int j = 0; j < 4; j++) { p[j] = lam[v][j] + c.logpi[j]; if (p[j] > mx) mx = p[j]; }
            } else {
                int x = Tg[(size_t)gpar[v] * len + g];
                if (c.miss[v] || gap[v] || x < 0 || x > 3) continue;
                /* log P(x -> j): row x of exp(log transposed) = column x */
                for (int j = 0; j < 4; j++) {
                    double pj = P[v][j * 4 + x];
                    p[j] = pj > 0 ? log(pj) + lam[v][j] : NEG_INF;
                    if (p[j] > mx) mx = p[j];
                }
            }
            for (int j = 0; j < 4; j++) p[j] = exp(p[j] - mx);
            Tg[(size_t)v * len + g] = (signed char)categorical(p, 4, slot_u(seed, (uint32_t)g, c2, c3, (uint32_t)v));
        }
    }
    free(lam); free(P);
    gt_ctx_free(&c);
}

/* Candidate scores of Sample_tree2: log P(branch's subtree | the new parent
 * sits on candidate t at substitution distance len) summed over the sites
 * g < len_sites where branch holds a base.  kind 0: t's sampled base x
 * (log of row x of P(len) times branch's partial), or the pi-average of that
 * message when t has no base; kind 1: the pi-average of the message (the
 * root candidate); kind 2: the root candidate's first-option score
 * log sum_i pi_i P(len)(b -> i), b = branch's own sampled base
 * (genetree.cpp:2221-2232), where an ambiguous leaf contributes nothing. */
static double gt_cand_score(const gt_ctx *c, const or_model *mm, const int *gch, const signed char *Tg, int len_sites,
                            const char *const *rows, long start, int gapchar, int branch, int t, int kind, double len,
                            double (*lam)[4], int *gap, int *tg, const double (*lam_branch)[4])
{
    double LP[16];
    if (len < 1e-8) log_identity(LP); else or_log_trans(mm, len, LP);
    double tot = 0;
    (void)gch; (void)rows; (void)start; (void)gapchar; (void)lam; (void)gap; (void)tg;
    for (int g = 0; g < len_sites; g++) {
        int xb = Tg[(size_t)branch * len_sites + g];
        if (xb >= 4) continue;
        if (kind == 2) {
            if (xb < 0) continue;
            double y[4];
            for (int i = 0; i < 4; i++) y[i] = LP[i * 4 + xb] + c->logpi[i];
            tot += log_exp_sum(y, 4);
            continue;
        }
        double M[4];
        log_multi(LP, lam_branch[g], M);
        int xt = kind == 0 ? Tg[(size_t)t * len_sites + g] : 4;
        if (kind == 0 && !c->miss[t] && xt >= 0 && xt < 4) tot += M[xt];
        else {
            double y[4];
            for (int j = 0; j < 4; j++) y[j] = M[j] + c->logpi[j];
            tot += log_exp_sum(y, 4);
        }
    }
    return tot;
}

int or_test_sample_tree(const or_model *m, double piA, const or_sptree *sp, int *gch, int *gpar, double *gh,
                        int *groot_io, const int *Z, double c_rate, double n_rate, double c2m, double n2m,
                        int root_ingrp, int branch, int indicator, const int *lens, const signed char *Tg,
                        double *loglik, double *logpz, uint64_t seed, int m_sweep, int slot, const char *const *rows,
                        long start, int GG, int gapchar, double gap_prop, int exact, int *cand_node,
                        double *cand_score, int *n_cand);

/* GTree::Sample_tree2 with Remove_branch and Graft (src/PhyloAcc-GT/
 * genetree.cpp:2138-2518, 737-923) on a gene tree held as topology +
 * heights, the unconstrained candidate set (BPP::cons_sample false: no
 * deep-coalescence species and the default branch-sampling threshold).
 *   Prune gene lineage `branch` with its parent bp: the sibling takes bp's
 *   place under bpp.  Walk up the species from branch's species ss: every
 *   gene node born in a species on the way (above branch in ss) is a
 *   candidate new grandparent, at substitution distance cur_len + rate *
 *   (h(t) - max(bottom, h(branch))), cur_len summing the species crossed
 *   below; in the species root a candidate above the gene root (N) at the
 *   fixed height curN = 10 theta_root / 2 (h(root) + 5 theta_root / 2 when
 *   the root is higher); stop after root_ingrp once a candidate exists.
 *   Draw a candidate with probability proportional to exp(score), then
 *   the new sibling (one of its two children, or the gene root for N) and
 *   bp's new height uniformly between where both lineages first share a
 *   species and the candidate (curN rule for N).  Graft bp there, then
 *   Metropolis-Hastings on the prefix likelihood, missing-gene term, the
 *   coalescent prior, the two uniform heights' densities and the candidate
 *   probabilities of the new and the old grandparent.
 * Tg [N * Tlen] is the history from the last GTree::Update_Tg draw, Tlen =
 * lens[indicator].  Draws: tag GT_TREE, sub-stream slot, sweep m: candidate
 * 0, child 1 (u < 1/2: the first), height 2 (floored at 1e-5), decision 3.
 * Returns 1 accepted (tree, loglik, logpz updated), 0 rejected (tree
 * restored); an old grandparent that is not a candidate (possible only
 * with root_ingrp below the species root) rejects the move. */
int or_gt_sample_tree(const or_model *m, double piA, const or_sptree *sp, int *gch, int *gpar, double *gh,
                      int *groot_io, const int *Z, double c_rate, double n_rate, double c2m, double n2m, int root_ingrp,
                      int branch, int indicator, const int *lens, const signed char *Tg, double *loglik,
                      double *logpz, uint64_t seed, int m_sweep, int slot, const char *const *rows, long start, int GG,
                      int gapchar, double gap_prop)
{
    return or_test_sample_tree(m, piA, sp, gch, gpar, gh, groot_io, Z, c_rate, n_rate, c2m, n2m, root_ingrp, branch,
                               indicator, lens, Tg, loglik, logpz, seed, m_sweep, slot, rows, start, GG, gapchar,
                               gap_prop, 0, NULL, NULL, NULL);
}

/* Test hook: or_gt_sample_tree with, when exact is set, the probabilities of
 * the new sibling's choice (1/2 among a candidate's two children, 1 for the
 * root candidate) added to the Hastings ratio for both directions -- the
 * reference leaves them out, so only with them is the move reversible with
 * respect to the coalescent prior.  cand_node / cand_score (size N + 2, may
 * be NULL) receive the candidates and their scores; the count is returned in
 * *n_cand. */
int or_test_sample_tree(const or_model *m, double piA, const or_sptree *sp, int *gch, int *gpar, double *gh,
                        int *groot_io, const int *Z, double c_rate, double n_rate, double c2m, double n2m,
                        int root_ingrp, int branch, int indicator, const int *lens, const signed char *Tg,
                        double *loglik, double *logpz, uint64_t seed, int m_sweep, int slot, const char *const *rows,
                        long start, int GG, int gapchar, double gap_prop, int exact, int *cand_node,
                        double *cand_score, int *n_cand)
{
    int S = sp->S, N = sp->N, groot = *groot_io;
    const double *H = sp->height;
    or_model mm;
    if (piA > 0) gt_model_at(m, piA, &mm); else mm = *m;
    int save_ch[2 * OR_MAXN], save_par[OR_MAXN];
    double save_h[OR_MAXN];
    memcpy(save_ch, gch, sizeof(int) * 2 * N);
    memcpy(save_par, gpar, sizeof(int) * N);
    memcpy(save_h, gh, sizeof(double) * N);
    double curG = or_gt_prior(sp, gch, gpar, gh, groot);
    int ss = or_gt_species(sp, gch, gh, branch);
    int bp = gpar[branch], cb = gch[2 * bp] == branch ? 0 : 1, bs = gch[2 * bp + 1 - cb], bpp = gpar[bp];
    double thr = sp->theta[N - 1];
    double curN = 10 * thr / 2;
    if (curN < gh[groot]) curN = gh[groot] + 5 * thr / 2;
    double curlow = 0;
    for (int x = ss; x < N; x = sp->parent[x])
        if (gt_present(sp, gch, gpar, gh, bs, x)) { curlow = fmax(H[x], fmax(gh[branch], gh[bs])); break; }
    double curupp = bpp < N ? gh[bpp] : curN;
    double cur_unif = -log(curupp - curlow);
    /* the branch's partials over the Tlen sites (current tree) */
    int Tlen = lens[indicator];
    gt_ctx c0;
    gt_ctx_init(&c0, &mm, sp, gch, gpar, gh, groot, Z, c_rate, n_rate, rows, start, GG, gapchar, gap_prop);
    double (*lam)[4] = (double (*)[4])malloc(sizeof(double) * 4 * N);
    double (*lb)[4] = (double (*)[4])malloc(sizeof(double) * 4 * (Tlen > 0 ? Tlen : 1));
    int gap[OR_MAXN], tg[OR_MAXN];
    for (int g = 0; g < Tlen; g++) {
        gt_site_partials(&c0, gch, rows, start, g, gapchar, lam, gap, tg);
        memcpy(lb[g], lam[branch], sizeof lb[g]);
    }
    /* Remove_branch */
    int cbp = bpp < N ? (gch[2 * bpp] == bp ? 0 : 1) : -1;
    gpar[bs] = bpp;
    if (bpp < N) gch[2 * bpp + cbp] = bs; else groot = bs;
    gpar[bp] = -1;
    /* candidates on the pruned tree */
    int cn[OR_MAXN + 2], csp[OR_MAXN + 2], ck[OR_MAXN + 2], nc = 0;
    double cl[OR_MAXN + 2], cs[OR_MAXN + 2];
    double cur_len = 0;
    for (int x = ss; x < N;) {
        double rate = Z[x] == 2 ? n_rate : (Z[x] == 1 ? c_rate : 1.0);
        double base = fmax(H[x], gh[branch]);
        int born[OR_MAXN], nb = 0;
        for (int v = S; v < N; v++) {
            if (v == bp || gch[2 * v] == -1) continue;
            if (v != groot && gpar[v] < 0) continue;
            if (or_gt_species(sp, gch, gh, v) == x) born[nb++] = v;
        }
        for (int i = 1; i < nb; i++)
            for (int j = i; j > 0 && gh[born[j - 1]] > gh[born[j]]; j--) { int t = born[j]; born[j] = born[j - 1]; born[j - 1] = t; }
        int up = sp->parent[x];
        if (nb == 0) {
            if (up == N) {
                cn[nc] = N; csp[nc] = x; cl[nc] = rate * (curN - base) + cur_len;
                ck[nc] = nc == 0 ? -1 : 2;
                nc++;
                break;
            }
            cur_len += (H[up] - base) * rate;
            x = up;
            continue;
        }
        for (int i = 0; i < nb; i++) {
            int t = born[i];
            if (x == ss && gh[t] <= gh[branch]) continue;
            cn[nc] = t; csp[nc] = x; cl[nc] = rate * (gh[t] - base) + cur_len; ck[nc] = 0;
            nc++;
        }
        if (up == N) {
            double ht = nc > 0 ? gh[cn[nc - 1]] : base;
            cn[nc] = N; csp[nc] = x; cl[nc] = cur_len + rate * (curN - ht); ck[nc] = 1;
            nc++;
        }
        if (up < N) cur_len += (H[up] - base) * rate;
        if (x == root_ingrp && nc > 0) break;
        x = up;
    }
    for (int i = 0; i < nc; i++)
        cs[i] = ck[i] < 0 ? 0.0 : gt_cand_score(&c0, &mm, gch, Tg, Tlen, rows, start, gapchar, branch, cn[i], ck[i],
                                                cl[i], lam, gap, tg, (const double (*)[4])lb);
    if (cand_node) for (int i = 0; i < nc; i++) { cand_node[i] = cn[i]; cand_score[i] = cs[i]; }
    if (n_cand) *n_cand = nc;
    double sp_p[OR_MAXN + 2];
    uint32_t c3 = tagword(0, OR_TAG_GT_TREE), c2 = (uint32_t)m_sweep;
    int get = 0;
    if (nc == 1) sp_p[0] = 1;
    else {
        double mx = NEG_INF, tot = 0;
        for (int i = 0; i < nc; i++) if (cs[i] > mx) mx = cs[i];
        for (int i = 0; i < nc; i++) { sp_p[i] = exp(cs[i] - mx); tot += sp_p[i]; }
        for (int i = 0; i < nc; i++) sp_p[i] /= tot;
        get = categorical(sp_p, nc, slot_u(seed, (uint32_t)slot, c2, c3, 0));
    }
    int nbpp = cn[get], nsib;
    double upp;
    if (nbpp == N) {
        nsib = groot;
        upp = 10 * thr / 2 > gh[nsib] ? 10 * thr / 2 : gh[nsib] + 5 * thr / 2;
    } else {
        int a = slot_u(seed, (uint32_t)slot, c2, c3, 1) < 0.5 ? 0 : 1;
        nsib = gch[2 * nbpp + a];
        upp = gh[nbpp];
    }
    double low = 0;
    for (int x = ss; x != sp->parent[csp[get]]; x = sp->parent[x])
        if (gt_present(sp, gch, gpar, gh, nsib, x)) { low = fmax(gh[branch], fmax(gh[nsib], H[x])); break; }
    double u = slot_u(seed, (uint32_t)slot, c2, c3, 2);
    if (u < 1e-5) u = 1e-5;
    double coal = u * (upp - low) + low;
    double new_unif = -log(upp - low);
    /* Graft */
    int it2p = gpar[nsib], cs_sib = it2p < N ? (gch[2 * it2p] == nsib ? 0 : 1) : -1;
    gh[bp] = coal;
    gpar[bp] = it2p;
    gpar[nsib] = bp;
    gch[2 * bp + cb] = branch;
    gch[2 * bp + 1 - cb] = nsib;
    if (it2p < N) gch[2 * it2p + cs_sib] = bp; else groot = bp;
    /* MH */
    double newG = or_gt_prior(sp, gch, gpar, gh, groot);
    double r = slot_u(seed, (uint32_t)slot, c2, c3, 3);
    gt_ctx c1;
    gt_ctx_init(&c1, &mm, sp, gch, gpar, gh, groot, Z, c_rate, n_rate, rows, start, GG, gapchar, gap_prop);
    double lnew[2] = {0, 0};
    for (int g = 0; g < Tlen; g++) {
        gt_site_partials(&c1, gch, rows, start, g, gapchar, lam, gap, tg);
        double y[4];
        for (int j = 0; j < 4; j++) y[j] = lam[groot][j] + c1.logpi[j];
        double t = log_exp_sum(y, 4);
        lnew[0] += t;
        if (g < lens[1]) lnew[1] += t;
    }
    int gm[OR_MAXN];
    for (int v = 0; v < N; v++) gm[v] = c1.miss[v];
    double lpz_new = or_gt_logpz(sp, gch, gh, gm, Z, c2m, n2m);
    gt_ctx_free(&c0); gt_ctx_free(&c1);
    free(lam); free(lb);
    int pos = -1;
    for (int i = 0; i < nc; i++) if (cn[i] == bpp) { pos = i; break; }
    int ok = 0;
    if (pos >= 0) {
        double d = lnew[indicator] - loglik[indicator];
        if (logpz[indicator] != 0) d += lpz_new - logpz[indicator];
        d += newG - curG + cur_unif - new_unif + log(sp_p[pos]) - log(sp_p[get]);
        if (exact) d += (bpp < N ? log(0.5) : 0.0) - (nbpp < N ? log(0.5) : 0.0);
        ok = log(r) < d;
    }
    if (ok) {
        loglik[0] = lnew[0];
        loglik[1] = lnew[1];
        if (logpz[1] == 0) logpz[0] = lpz_new;
        else logpz[0] = logpz[1] = lpz_new;
        *groot_io = groot;
        return 1;
    }
    memcpy(gch, save_ch, sizeof(int) * 2 * N);
    memcpy(gpar, save_par, sizeof(int) * N);
    memcpy(gh, save_h, sizeof(double) * N);
    return 0;
}
