/*
 * jess_oracle.h -- CPU restatement of Jess template matching (TEST INFRASTRUCTURE).
 *
 * This library is the parity oracle and the CPU baseline of the test-suite and
 * benchmark. It is never linked into, loaded by or called from the product
 * library (pyjess_amd/csrc). Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it, and only as the checker.
 *
 * It restates, in plain C, the algorithm of pyjess v0.10.1 + the patched Jess
 * engine (reference files cited per function in jess_oracle.c).
 */
#ifndef JESS_ORACLE_H
#define JESS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Molecule in structure-of-arrays form. Names are 4 bytes per atom with
 * '_' in place of blanks (pyjess `encode_token`, _jess.pyx:133-140); residue
 * names are 3 bytes + NUL per atom (Atom.h resName[4]). */
typedef struct {
    int32_t n;
    const double *xyz;     /* 3*n */
    const char *name;      /* 4*n */
    const char *resname;   /* 4*n */
    const char *chain2;    /* n   (chainID2, the only chain byte check() uses) */
    const int32_t *resseq; /* n */
    const int32_t *serial; /* n */
} jo_mol;

/* A set of templates, flattened. Template t owns atoms [atom_off[t], atom_off[t+1]). */
typedef struct {
    int32_t n_templates;
    const int32_t *atom_off;    /* T+1 */
    const int32_t *mode;        /* A  (TessAtom.code) */
    const int32_t *resseq;      /* A */
    const char *chain2;         /* A */
    const double *pos;          /* 3*A */
    const double *weight;       /* A  (distWeight) */
    const int32_t *name_off;    /* A+1, index into names (4 bytes each) */
    const char *names;
    const int32_t *resname_off; /* A+1, index into resnames (4 bytes each) */
    const char *resnames;
} jo_tpls;

typedef struct {
    double distance_cutoff;       /* t */
    double max_dynamic_distance;  /* s */
    int32_t ignore_chain;         /* 0 none, 1 residues, 2 atoms (IgnoreType) */
    int32_t reorder;
    int32_t kd_right_first;       /* child push order of the kd DFS */
} jo_params;

/* Ordered stream of every complete match of a query, templates in scan order. */
typedef struct jo_stream jo_stream;

/* Error codes */
#define JO_OK 0
#define JO_ENOMEM (-1)
#define JO_EMODE (-2)   /* a match mode whose semantics are not pinned */

int32_t jo_query(const jo_mol *mol, const jo_tpls *tpls, const jo_params *p,
                 const int32_t *scan_order, int32_t n_scan, jo_stream **out);
int64_t jo_stream_count(const jo_stream *s);
int64_t jo_stream_atom_total(const jo_stream *s);
void jo_stream_copy(const jo_stream *s, int32_t *tpl, int64_t *atom_off, int32_t *atoms,
                    double *rmsd, double *rot, double *c0, double *c1);
void jo_stream_free(jo_stream *s);

/* One superposition (n point pairs, molecule points a, template points b). */
void jo_superpose(const double *a, const double *b, int32_t n,
                  double *rmsd, double *rot, double *c0, double *c1);

/* Brute-force enumeration of complete matches of one template (tiny inputs only):
 * every n-tuple of candidates passing all pairwise annuli and check(), in
 * lexicographic order of candidate-list positions. Returns count; writes up to
 * `cap` tuples (template order) into out. */
int64_t jo_bruteforce(const jo_mol *mol, const jo_tpls *tpls, int32_t t, const jo_params *p,
                      int32_t *out, int64_t cap);

/* Candidate list of template atom `k` of template `t` (TessTemplate_candidates).
 * Returns the count (or JO_EMODE), writes up to `cap` atom indices. */
int32_t jo_candidates(const jo_mol *mol, const jo_tpls *tpls, int32_t t, int32_t k,
                      int32_t *out, int32_t cap);

/* kd leaf order of a candidate set (positions into the set list). */
int32_t jo_kd_order(const jo_mol *mol, const int32_t *set_atoms, int32_t n, int32_t *out);

/* Algorithmic work of a query (see jess_oracle.c). out: 4 counters. */
int32_t jo_work(const jo_mol *mol, const jo_tpls *tpls, const jo_params *p, int64_t *out);

/* Multi-threaded batch: every (template, molecule) pair; per-pair counts of
 * complete matches and of hits (rmsd <= threshold, finite rotation), and an
 * order-independent 64-bit digest of the hit set. Used as the CPU baseline. */
int32_t jo_batch(const jo_mol *mols, int32_t n_mols, const jo_tpls *tpls, const jo_params *p,
                 double rmsd_threshold, int32_t n_threads,
                 int64_t *cm_count, int64_t *hit_count, uint64_t *digest);

#ifdef __cplusplus
}
#endif
#endif
