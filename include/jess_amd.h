/*
 * jess_amd.h -- C ABI of the MI355X (gfx950) Jess template-matching library.
 *
 * Plain pointers and sizes only. All functions return 0 on success or a
 * negative JAM_E* code; jam_last_error() then describes the failure (the
 * message is per thread). Nothing here aborts the process.
 *
 * Ownership: every object created by a jam_*_create function is owned by the
 * caller and released with the matching jam_*_free. Input arrays are copied
 * during the call; the library keeps no pointer to caller memory.
 *
 * Each declaration names the reference interface (althonos/pyjess v0.10.1)
 * that it replaces; paths are relative to the reference checkout.
 */
#ifndef JESS_AMD_H
#define JESS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JAM_OK 0
#define JAM_ENOMEM (-1)     /* host or device allocation failed */
#define JAM_EINVAL (-2)     /* invalid argument */
#define JAM_EMODE (-3)      /* template uses a match mode other than 0, 1 and 3 */
#define JAM_EDEVICE (-4)    /* HIP runtime error or no usable gfx950 device */
#define JAM_ETOOBIG (-5)    /* template larger than JAM_MAX_TEMPLATE_ATOMS */
#define JAM_ESTALE (-6)     /* a result's events were overwritten by a later query */

#define JAM_MAX_TEMPLATE_ATOMS 32

/* ignore_chain values (IgnoreType, include/jess/template.pxd:5-8) */
#define JAM_IGNORE_NONE 0
#define JAM_IGNORE_RESIDUES 1
#define JAM_IGNORE_ATOMS 2

typedef struct jam_context jam_context;
typedef struct jam_templates jam_templates;
typedef struct jam_molecules jam_molecules;
typedef struct jam_result jam_result;

/* One molecule, structure of arrays. Names are 4 bytes per atom with '_' for
 * blanks (pyjess encode_token, src/pyjess/_jess.pyx:133-140); residue names
 * are 3 bytes + NUL per atom (Atom.resName, include/jess/atom.pxd:3-19). */
typedef struct {
    int32_t n_atoms;
    const double *xyz;      /* 3 * n_atoms, x y z per atom */
    const char *names;      /* 4 * n_atoms */
    const char *resnames;   /* 4 * n_atoms */
    const char *chain2;     /* n_atoms: second chain byte (chainID2) */
    const int32_t *resseq;  /* n_atoms */
    const int32_t *serial;  /* n_atoms */
} jam_molecule_desc;

/* A list of templates, flattened; template t owns atoms
 * [atom_off[t], atom_off[t+1]). Per atom: the TessAtom fields
 * (include/jess/tess_atom.pxd:3-17). names[4*k] for k in
 * [name_off[a], name_off[a+1]) are the atom names of atom a, resnames
 * likewise (3 bytes + NUL each). */
typedef struct {
    int32_t n_templates;
    const int32_t *atom_off;     /* n_templates + 1 */
    const int32_t *mode;         /* per atom: match mode */
    const int32_t *resseq;
    const char *chain2;
    const double *pos;           /* 3 per atom */
    const double *weight;        /* distance weight */
    const int32_t *name_off;     /* n_atoms + 1 */
    const char *names;
    const int32_t *resname_off;  /* n_atoms + 1 */
    const char *resnames;
} jam_template_desc;

typedef struct {
    double rmsd_threshold;        /* Query.rmsd_threshold */
    double distance_cutoff;       /* t of Jess_query */
    double max_dynamic_distance;  /* s of Jess_query */
    int32_t ignore_chain;         /* JAM_IGNORE_* */
    int32_t reorder;              /* 0 or 1 */
} jam_params;

/* ---- device context ------------------------------------------------- */

/* Opens device `device` (HIP ordinal), checks it is gfx950, creates one
 * stream and a growable workspace. No reference counterpart: pyjess is
 * CPU-only; this is the device state behind every query. */
int32_t jam_context_create(int32_t device, jam_context **out);
void jam_context_free(jam_context *ctx);

/* ---- templates -------------------------------------------------------- */

/* Replaces Jess_create + Jess_addTemplate (include/jess/jess.pxd:9-11,
 * patches/Jess.c.patch:63-72) and the distance matrix of Template.__init__
 * (src/pyjess/_jess.pyx:1895-1901, TessAtom_distance patches/TessAtom.c.patch:229-245).
 * Copies the templates to the device, deduplicates atom types across them
 * and computes the pairwise template distances. Templates keep the index
 * they have in `desc`; queries scan them in reverse order, as Jess does. */
int32_t jam_templates_create(jam_context *ctx, const jam_template_desc *desc, jam_templates **out);
void jam_templates_free(jam_templates *t);

/* ---- molecules -------------------------------------------------------- */

/* Replaces Molecule.__init__ + ResIndex_create (src/pyjess/_jess.pyx:632-684,
 * patches/ResIndex.c.patch:34-119): copies `n` molecules to the device as one
 * batch and builds each residue-name index. */
int32_t jam_molecules_create(jam_context *ctx, const jam_molecule_desc *mols, int32_t n,
                             jam_molecules **out);
void jam_molecules_free(jam_molecules *m);

/* flags of jam_molecules_parse_pdb (PDBMoleculeParser's skip_hetatm and
 * ignore_endmdl, src/pyjess/_io.py:52-66) */
#define JAM_PDB_SKIP_HETATM 1
#define JAM_PDB_IGNORE_ENDMDL 2

/* Replaces PDBMoleculeParser.load (src/pyjess/_io.py:68-103), Atom.loads for
 * every ATOM and HETATM record, Molecule.__init__ and ResIndex_create
 * (src/pyjess/_jess.pyx:632-684, patches/ResIndex.c.patch:34-119) for n PDB
 * texts at once, on the device: text k is the `lengths[k]` bytes at
 * `texts[k]` (host memory, copied to the device as it is) and becomes
 * molecule k of the batch. Lines end at '\n', '\r\n' or '\r'; ATOM records
 * (and HETATM records unless JAM_PDB_SKIP_HETATM) up to the first ENDMDL
 * record (all of them with JAM_PDB_IGNORE_ENDMDL) are read in fixed columns,
 * numbers parsed exactly as Python's int() and float() parse decimal text.
 * JAM_EINVAL, naming the text and line, for an mmCIF data tag, a numeric
 * field that does not parse (or that uses a form other than decimal digits
 * with an optional point and exponent), or a byte above 0x7F in the first 80
 * columns of a record. HEADER records are not read. The context keeps the
 * parse's temporary device buffers for its next parse. */
int32_t jam_molecules_parse_pdb(jam_context *ctx, const char *const *texts, const int64_t *lengths, int32_t n,
                                int32_t flags, jam_molecules **out);

/* Every atom field, host buffers of one entry per atom of the batch (any may
 * be NULL): names, resnames, segid 4 bytes, chains (chainID1, chainID2) and
 * element 2 bytes per atom, in the encodings of jam_molecule_desc. */
typedef struct {
    double *xyz;
    char *names;
    char *resnames;
    char *chains;
    int32_t *resseq;
    int32_t *serial;
    char *altloc;
    char *icode;
    float *occupancy;
    float *tempfactor;
    char *segid;
    char *element;
    int32_t *charge;
} jam_atom_fields;

/* Atom offsets of the molecules of a batch (n + 1 entries) and the number
 * of residue-name buckets (either may be NULL). No reference counterpart. */
int32_t jam_molecules_sizes(const jam_molecules *m, int64_t *atom_off, int64_t *n_buckets);
/* Copies the atom fields of a batch to the host (Molecule's arrays,
 * src/pyjess/_jess.pyx:243-907); fields other than xyz, names, resnames and
 * resseq exist only for a batch made by jam_molecules_parse_pdb. */
int32_t jam_molecules_copy_atoms(const jam_molecules *m, const jam_atom_fields *fields);
/* Copies the residue-name index of a batch (ResIndex, patches/ResIndex.c.patch:34-119):
 * perm (per molecule, its atoms in (upper-cased residue name, serial)
 * order, molecule-local indices), res_off (n + 1: each molecule's buckets),
 * and per bucket its key (upper-cased packed name) and range in perm. */
int32_t jam_molecules_residue_index(const jam_molecules *m, int32_t *perm, int32_t *res_off, uint32_t *res_key,
                                    int32_t *res_begin, int32_t *res_end);

/* ---- queries ---------------------------------------------------------- */

/* Replaces Jess_query + the JessQuery_next loop + JessQuery_superposition +
 * Superposition_rmsd/rotation/centroid (include/jess/jess.pxd:11-21,
 * include/jess/super.pxd:1-14, patches/Jess.c.patch:75-216) for every
 * (template, molecule) pair of the batch at once. The result holds every
 * complete match whose RMSD is <= rmsd_threshold ("events": hits, and the
 * matches whose rotation has a NaN, which pyjess skips with a warning), each
 * with its pair, its rank in the reference enumeration order of that pair,
 * the matched atoms (molecule-local indices, template order), RMSD, rotation
 * and centroids. The events are ordered as pyjess reports them: by molecule,
 * then template in scan order, then rank; Query.__next__'s policies
 * (best_match, max_candidates) are applied by the caller in one pass over
 * them. They stay in device memory, owned by the context, until the next
 * jam_query on that context: jam_result_copy fetches them before then. */
int32_t jam_query(jam_context *ctx, const jam_templates *t, const jam_molecules *m,
                  const jam_params *params, jam_result **out);

/* ---- results ---------------------------------------------------------- */

int64_t jam_result_n_events(const jam_result *r);
int32_t jam_result_stride(const jam_result *r);   /* atoms per event row */
int32_t jam_result_n_pairs(const jam_result *r);  /* n_molecules * n_templates */
/* Copies the events from device memory into the caller's host buffers
 * (n_events rows each). Any pointer may be NULL to skip that field.
 * pair = molecule * n_templates + template. flag: 1 hit, 2 NaN rotation.
 * JAM_ESTALE once another jam_query has run on the result's context; the
 * result must not outlive its context. */
int32_t jam_result_copy(const jam_result *r, int32_t *pair, int64_t *rank, int32_t *flag,
                        double *rmsd, double *rotation9, double *c0_3, double *c1_3,
                        int32_t *atoms_stride);
/* Number of complete matches (before the RMSD threshold) of every pair. */
int32_t jam_result_pair_matches(const jam_result *r, int64_t *counts);
/* Measurement (no reference counterpart): per pair, the number of hits
 * (events with flag 1) and, summed mod 2^64 over them, a 64-bit digest of
 * each hit -- h = mix64(template << 32 ^ (mol_base + molecule)), then
 * h = mix64(h ^ atom) for each matched atom in template order, mix64 being
 * SplitMix64's finaliser -- an order-independent digest of the hit set that
 * a CPU run computes the same way. Computed on the device from the events in
 * HBM (n_pairs values each); JAM_ESTALE once a later query has run. */
int32_t jam_result_pair_digests(const jam_result *r, int32_t mol_base, int32_t n_templates, int64_t *hit_counts,
                                uint64_t *digests);
/* Kernel durations of the last query, milliseconds: [candidates, kd order,
 * join, superposition, total device time]. */
int32_t jam_result_timings(const jam_result *r, double *ms5);
/* Summed duration (milliseconds, HIP events on the library's stream) of the
 * join's k_join launches of the last query -- one per join level, the inner
 * loop of Scanner_next (patches/Scanner.c.patch:232-334)
 * -- and the number of those launches; -1 ms and 0 launches unless the
 * context records them (jam_context_set_profiling). */
int32_t jam_result_join_kernels(const jam_result *r, double *ms, int32_t *launches);
/* Partial matches expanded at each join level L = 1 .. 31 of the last query,
 * written to counts32[L] (32 entries; counts32[0] is unused): the number of open Scanner_next
 * states per depth (patches/Scanner.c.patch:232-334) summed over pairs. */
int32_t jam_result_level_partials(const jam_result *r, int64_t *counts32);
/* Switches the per-launch events of the join kernels on (1) or off (0, the
 * default): each event record costs a few microseconds of stream time. No
 * counterpart in the reference (measurement only). */
int32_t jam_context_set_profiling(jam_context *ctx, int32_t on);
void jam_result_free(jam_result *r);

/* Events kept in device memory of their own: one device-to-device copy of a
 * result's events (queued on the context's stream, before any later query),
 * so that a caller can copy some fields now (pair, rank, flag, rmsd: what
 * Query.__next__'s policy needs, src/pyjess/_jess.pyx:2238-2276) and the
 * others (rotation, centroids, atoms: what a Hit's transformation and atoms
 * read, _jess.pyx:2343-2482) only if and when they are read. Same arguments
 * and layout as jam_result_copy; they stay valid after later queries on the
 * context, and hold their device memory until jam_events_free. */
typedef struct jam_events jam_events;
int32_t jam_result_keep(const jam_result *r, jam_events **out);
int32_t jam_events_copy(const jam_events *e, int32_t *pair, int64_t *rank, int32_t *flag, double *rmsd,
                        double *rotation9, double *c0_3, double *c1_3, int32_t *atoms_stride);
void jam_events_free(jam_events *e);

/* ---- standalone superposition ---------------------------------------- */

/* Batched least-squares superposition of `n_sets` point-set pairs of `k`
 * points each, on the device: a (molecule side) and b (template side) are
 * n_sets*k*3 doubles in host memory. Replaces Superposition_align +
 * Superposition_rmsd/rotation/centroid (include/jess/super.pxd:1-14) for a
 * batch. Outputs are n_sets, 9*n_sets, 3*n_sets, 3*n_sets doubles. */
int32_t jam_superpose_batch(jam_context *ctx, const double *a, const double *b, int32_t k,
                            int64_t n_sets, double *rmsd, double *rotation9, double *c0_3,
                            double *c1_3);

/* Size of the per-device scratch array of jam_superpose_batch_noalloc. */
#define JAM_NOALLOC_SCRATCH_BYTES (32u << 20)

/* jam_superpose_batch without any device allocation, for callers that cannot
 * allocate device memory (serving loops and harnesses that forbid
 * hipMalloc). No counterpart in the reference. Same arguments, kernels and
 * results as jam_superpose_batch, but inputs and outputs are staged in a
 * statically sized device array (JAM_NOALLOC_SCRATCH_BYTES per device, held
 * until the process ends) instead of context buffers: batches needing more
 * than 8 * (6k + 16) bytes per set in all (about 32 MiB) return
 * JAM_ETOOBIG. The array is shared by every context of a device and keeps
 * the last call's contents (nothing in it needs to start at zero); calls on
 * one device from one process run one at a time (a lock), calls from
 * several processes on one device would race. */
int32_t jam_superpose_batch_noalloc(jam_context *ctx, const double *a, const double *b, int32_t k,
                                    int64_t n_sets, double *rmsd, double *rotation9, double *c0_3,
                                    double *c1_3);

/* ---- standalone kd-tree leaf order ----------------------------------- */

/* Replaces KdTree_create / KdTreeNode_create (patches/KdTree.c.patch:130-185,
 * 348-420, with qselect_r, patches/qselect.h.patch:50-116) for a batch of
 * point sets: the leaf order, left subtree first, of the kd-tree the
 * reference builds over each set (median split at n/2 on axis depth % 3 by
 * median-of-three QuickSelect). xyz: 3 doubles per point, host memory; set s
 * owns points [off[s], off[s+1]) (off[0] == 0, never decreasing);
 * order[off[s] + i] receives the set-local index of the point at leaf i. The
 * same device passes order every candidate set of jam_query. */
int32_t jam_kd_order(jam_context *ctx, const double *xyz, const int64_t *off, int32_t n_sets, int32_t *order);

/* ---- differentiable superposition RMSD --------------------------------- */

/* RMSD of the least-squares superposition of a onto b for a batch of point
 * sets of any sizes, all arrays in DEVICE memory: a and b hold n_points
 * points (3 doubles each, x y z); set s owns points [off[s], off[s+1]), with
 * off[0] == 0, off never decreasing and off[n_sets] == n_points. rmsd
 * receives n_sets values, rot9 (may be NULL) the 9 row-major entries of each
 * optimal proper rotation. The quantity of Superposition_rmsd/rotation
 * (include/jess/super.pxd:1-14), computed for differentiation: rotation from
 * a one-sided Jacobi SVD of the correlation, RMSD from the residuals, so that
 * near-degenerate sets keep full precision (a set of rank < 2 gets one of its
 * optimal rotations instead of NaN). Asynchronous on `stream` (a hipStream_t;
 * NULL for the default stream): the offsets are not read on the host, so the
 * call cannot reject malformed ones; the kernels clamp every set's range into
 * [0, n_points), so malformed offsets give meaningless values but never an
 * access outside a and b. */
int32_t jam_rmsd_forward(const double *a, const double *b, const int64_t *off, int64_t n_sets, int64_t n_points,
                         double *rmsd, double *rot9, void *stream);
/* Gradient of sum_s grad_rmsd[s] * rmsd[s] with respect to a and b (device
 * arrays laid out like a and b; either may be NULL), given the rmsd of
 * jam_rmsd_forward and, optionally, its rotations (rot9; NULL recomputes
 * them). Every point of a set with at least one point is written, so with
 * valid offsets every row of grad_a and grad_b is. No counterpart in the
 * reference, which has no gradients:
 * d rmsd / d a_i = (A_i - R^T B_i) / (n rmsd) and
 * d rmsd / d b_i = (B_i - R A_i) / (n rmsd), A and B the centred points, R
 * the optimal proper rotation; zero where rmsd is 0. Asynchronous on
 * `stream`, offsets clamped as in jam_rmsd_forward. */
int32_t jam_rmsd_backward(const double *a, const double *b, const int64_t *off, int64_t n_sets, int64_t n_points,
                          const double *rmsd, const double *rot9, const double *grad_rmsd, double *grad_a,
                          double *grad_b, void *stream);

/* ---- misc ------------------------------------------------------------- */

/* Waits until all work queued on the context's device has finished
 * (hipDeviceSynchronize). No reference counterpart; used to bracket timings. */
int32_t jam_synchronize(jam_context *ctx);

const char *jam_last_error(void);
/* Library version string. */
const char *jam_version(void);
/* 1 when a usable gfx950 device is visible, else 0 (never initialises more than needed). */
int32_t jam_device_available(void);

#ifdef __cplusplus
}
#endif
#endif
