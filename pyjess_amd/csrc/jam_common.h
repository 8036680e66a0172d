// jam_common.h -- device-side layouts shared by the kernels and the host API.
//
// HBM layout (all arrays device resident, structure of arrays):
//
//   molecules   concatenated atoms of every molecule of a batch; molecule m owns
//               atoms [atom_off[m], atom_off[m+1]); per atom xyz (AoS, 24 B),
//               raw 4-byte name and residue name, chainID2, resSeq; plus the
//               residue-name index: perm (atoms sorted by upper-cased residue
//               name, then serial), bucket keys and bucket ranges.
//   templates   per template atom: atom-type id, chainID2, resSeq, position,
//               distance weight; per template an n*n distance matrix; per
//               atom type (deduplicated across templates): mode, names,
//               upper-cased residue names in template order.
//   candidate   one set per (molecule, atom type): the candidate list in
//   sets        reference order (molecule-local atom index), and the set in
//               kd-tree DFS leaf order (a 32-byte entry: x, y, z, chain,
//               resSeq; and the atom index) with one bounding box per 64
//               consecutive entries and one per 16.
//   join        frontier rows (pair, one entry index per placed level),
//               double buffered, and the complete matches.
#pragma once

#include <stdint.h>

#define JAM_WAVE 64
// entries per sub-block: the join tests a sub-block's box before its entries
#define JAM_SUB 16

struct DevMols {
    int n_mols;
    const int *atom_off;      // M+1
    const double *xyz;        // 3 per atom
    const uint32_t *name;     // raw packed name, byte 0 = first character
    const uint32_t *resname;  // raw packed residue name
    const int *chain;         // chainID2 (signed char value)
    const int *resseq;
    const int *perm;          // per molecule: local atom indices, residue index order
    const int *res_off;       // M+1: bucket range of molecule m
    const uint32_t *res_key;  // per bucket: upper-cased packed residue name
    const int *res_begin;     // per bucket: first position in perm (molecule local)
    const int *res_end;       // per bucket: one past the last position
};

struct DevTpls {
    int n_templates;
    int n_types;
    int stride;               // largest template (atoms per match row)
    const int *atom_off;      // T+1
    const int *atom_type;     // per atom
    const int *chain;         // per atom
    const int *resseq;        // per atom
    const double *pos;        // 3 per atom
    const double *weight;     // per atom
    const int *dist_off;      // T: offset of the n*n distance matrix
    const double *dist;
    const int *type_mode;     // U
    const int *type_name_off; // U+1
    const uint32_t *type_names;  // raw packed names
    const int *type_rn_off;   // U+1
    const uint32_t *type_rns;    // upper-cased packed residue names, template order
};

// One candidate-set entry as the join reads it: coordinates, chainID2 and
// resSeq of the atom, 32 bytes (two 16-byte loads).
struct SetEnt {
    double x, y, z;
    int ch, res;
};

struct DevSets {
    int n_sets;               // M * U
    int64_t cap;              // capacity of the entry arrays
    int64_t blk_cap;          // capacity of the block boxes
    int *cnt;                 // per set
    int64_t *off;             // per set: exclusive scan of cnt
    int *list;                // reference-order candidate list (molecule-local atom)
    int *kidx;                // kd position -> list position (set-local), per entry
    int *l2k;                 // list position -> kd position (set-local), per entry
    SetEnt *kent;             // kd DFS order
    int *katom;               // kd DFS order: molecule-local atom
    double *bbox;             // 6 per block of 64 entries: lo xyz, hi xyz
    double *sbox;             // 6 per sub-block of JAM_SUB entries
    int64_t *total;           // device scalar: sum of cnt
    int *maxcnt;              // device scalar: largest cnt
    unsigned int *kd_err;     // device word, 0 unless a kd pass hit its round bound
};

#define JAM_MAX_LEVELS 32

// What the join needs about level l of a pair: its candidate set (id, size,
// first entry) and the template atom placed at that level.
struct PairLevel {
    int set, cnt, off, k;
};

struct DevPairs {
    int n_pairs;
    const long long *n_items; // level-0 partial matches in all (device scalar)
    long long items_cap;      // capacity of item_pair
    unsigned char *ord;       // JAM_MAX_LEVELS per pair: level -> template atom
    PairLevel *lvl;           // JAM_MAX_LEVELS per pair
    int *nchunk;              // level-0 partial matches per pair
    int *n1;                  // first level-0 list index (1 for one-atom templates)
    long long *item_base;     // exclusive scan of nchunk
    int *item_pair;           // pair of every level-0 partial
};

// One level of the join. Rows are partial matches: the pair and, per placed
// level, the entry index of its atom (level 0: into the list entries,
// level l >= 1: into the kd entries). Rows are appended in any order; the
// complete matches are sorted into the reference order at the end.
struct DevJoin {
    int level;                          // L: the level being added
    int stride;                         // ints per row of the complete matches
    int in_stride;                      // ints per row of the input frontier: L (level L-1's rows)
    int out_stride;                     // ints per row of the output frontier: L + 1
    const unsigned long long *n_in;     // rows of the input frontier (may exceed in_cap)
    long long in_cap;
    const int *in_pair;
    const int *in_ent;
    unsigned long long *n_out;          // slots taken in the output frontier
    long long out_cap;
    int *out_pair, *out_ent;
    unsigned long long *n_done;         // complete matches
    long long done_cap;
    int *done_pair, *done_ent;
    unsigned int *overflow;             // set when a row did not fit
};

// The events of one query, packed field after field for exactly n events
// (the 8-byte fields first), so that copying them out is one transfer.
struct EventLayout {
    size_t rank, rmsd, rot, c0, c1, pair, flag, atoms, bytes;
};

__host__ __device__ inline EventLayout event_layout(int64_t n, int stride)
{
    EventLayout L;
    size_t o = 0;
    L.rank = o;
    o += sizeof(int64_t) * n;
    L.rmsd = o;
    o += sizeof(double) * n;
    L.rot = o;
    o += sizeof(double) * 9 * n;
    L.c0 = o;
    o += sizeof(double) * 3 * n;
    L.c1 = o;
    o += sizeof(double) * 3 * n;
    L.pair = o;
    o += sizeof(int32_t) * n;
    L.flag = o;
    o += sizeof(int32_t) * n;
    L.atoms = o;
    o += sizeof(int32_t) * (size_t)stride * n;
    L.bytes = o;
    return L;
}

struct DevEvents {
    int64_t cap;
    int stride;
    unsigned char *base;        // event_layout(count, stride)
    unsigned long long *count;  // device scalar
};

struct QueryParams {
    double rmsd_threshold;
    double distance_cutoff;
    double max_dynamic_distance;
    int ignore_chain;
    int reorder;
    int chunk;                // level-0 candidates per work item
};

__host__ __device__ static inline uint32_t jam_upper4(uint32_t v)
{
    uint32_t r = 0;
    for (int i = 0; i < 4; i++) {
        uint32_t c = (v >> (8 * i)) & 0xFFu;
        if (c >= 'a' && c <= 'z') c &= ~0x20u;
        r |= c << (8 * i);
    }
    return r;
}
