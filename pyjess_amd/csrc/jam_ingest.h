// jam_ingest.h -- PDB text to device molecules, on gfx950.
//
// The reference reads a PDB file line by line in Python (PDBMoleculeParser,
// src/pyjess/_io.py:68-103): ATOM records, and HETATM records unless skipped,
// become atoms (Atom.loads: fixed columns), up to the first ENDMDL record
// unless it is ignored; mmCIF data tags are an error. Molecule.__init__ then
// builds the residue-name index (ResIndex_create,
// patches/ResIndex.c.patch:34-119): atoms sorted by case-insensitive residue
// name, then serial number, grouped into one bucket per name.
//
// Here the texts of a batch of molecules (concatenated, each ending with a
// line feed) go through:
//   k_ing_count / k_ing_starts  line starts (after '\n', '\r\n' or a lone '\r',
//                               as Python's universal newlines read them),
//                               counted per block, scanned, written in order
//   k_ing_classify              record type of every line, text of every line,
//                               first ENDMDL and first mmCIF tag of every text
//   k_ing_select                atom records in order (a scan of 0/1 flags)
//   k_ing_parse                 thread per atom record: the fixed columns
//   k_ing_names ... k_ing_ranges the residue-name index: distinct names per
//                               molecule (LDS hash set, sorted), per-atom
//                               bucket, then one stable radix sort of the
//                               atoms by (bucket, serial) and the buckets'
//                               ranges
// Numbers are parsed exactly: a decimal field of at most 19 significant
// digits and a power of ten up to 22 is one correctly rounded division or
// multiplication of two exact doubles (Clinger's fast path), which is what
// Python's float() returns; other numeric forms are reported as errors.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jam_common.h"
#include "jam_wave.h"

namespace jam {

#define ING_BPT 16          // bytes per thread when splitting lines
#define ING_THREADS 256
#define ING_MAX_NAMES 2048  // distinct residue names per molecule

// line and record kinds
#define ING_OTHER 0
#define ING_ATOM 1
#define ING_HETATM 2
#define ING_ENDMDL 3
#define ING_MMCIF 4

// parse errors (bit flags of the per-atom error word)
#define ING_EBADFIELD 1     // a numeric field does not parse
#define ING_ENONASCII 2     // a byte above 0x7F in the first 80 columns

struct IngText {
    const unsigned char *buf;  // the texts, concatenated, each ending with '\n'
    long long n_bytes;
    const long long *text_off; // n_texts + 1
    int n_texts;
    int flags;                 // JAM_PDB_* of the ABI
};

__device__ static inline bool ing_line_start(const unsigned char *b, long long i)
{
    if (i == 0) return true;
    const unsigned char p = b[i - 1];
    return p == '\n' || (p == '\r' && b[i] != '\n');
}

// in-block exclusive scan of one value per thread (ING_THREADS threads)
__device__ static inline long long ing_block_scan(long long v, long long *lds, long long &total)
{
    const int t = threadIdx.x, lane = lane_id(), w = t / JAM_WAVE;
    long long x = v;
    for (int sh = 1; sh < JAM_WAVE; sh <<= 1) {
        const long long y = __shfl_up(x, sh, JAM_WAVE);
        if (lane >= sh) x += y;
    }
    if (lane == JAM_WAVE - 1) lds[w] = x;
    __syncthreads();
    long long base = 0;
    total = 0;
    for (int k = 0; k < ING_THREADS / JAM_WAVE; k++) {
        if (k < w) base += lds[k];
        total += lds[k];
    }
    __syncthreads();
    return base + x - v;
}

// Line starts of each block's ING_THREADS * ING_BPT bytes.
__global__ void __launch_bounds__(ING_THREADS) k_ing_count(IngText tx, long long *block_lines)
{
    __shared__ long long lds[ING_THREADS / JAM_WAVE];
    const long long b0 = ((long long)blockIdx.x * ING_THREADS + threadIdx.x) * ING_BPT;
    long long c = 0;
    for (int k = 0; k < ING_BPT; k++) {
        const long long i = b0 + k;
        if (i < tx.n_bytes && ing_line_start(tx.buf, i)) c++;
    }
    long long total;
    ing_block_scan(c, lds, total);
    if (threadIdx.x == 0) block_lines[blockIdx.x] = total;
}

// Writes the line starts in order (block_off: exclusive scan of block_lines).
__global__ void __launch_bounds__(ING_THREADS) k_ing_starts(IngText tx, const long long *block_off, long long *starts)
{
    __shared__ long long lds[ING_THREADS / JAM_WAVE];
    const long long b0 = ((long long)blockIdx.x * ING_THREADS + threadIdx.x) * ING_BPT;
    long long c = 0;
    for (int k = 0; k < ING_BPT; k++) {
        const long long i = b0 + k;
        if (i < tx.n_bytes && ing_line_start(tx.buf, i)) c++;
    }
    long long total;
    long long o = block_off[blockIdx.x] + ing_block_scan(c, lds, total);
    for (int k = 0; k < ING_BPT; k++) {
        const long long i = b0 + k;
        if (i < tx.n_bytes && ing_line_start(tx.buf, i)) starts[o++] = i;
    }
}

// end of line li (one past its last character, terminators excluded)
__device__ static inline long long ing_line_end(const IngText &tx, const long long *starts, long long n_lines,
                                                long long li)
{
    long long e = li + 1 < n_lines ? starts[li + 1] : tx.n_bytes;
    const long long s = starts[li];
    while (e > s && (tx.buf[e - 1] == '\n' || tx.buf[e - 1] == '\r')) e--;
    return e;
}

__device__ static inline int ing_text_of(const IngText &tx, long long pos)
{
    int lo = 0, hi = tx.n_texts;   // the last text whose start is <= pos
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tx.text_off[mid] <= pos) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ static inline unsigned char ing_lower(unsigned char c) { return (c >= 'A' && c <= 'Z') ? c + 32 : c; }

// Kind and text of every line; first ENDMDL and first mmCIF tag per text
// (first_endmdl, first_mmcif: n_texts words each, preset to n_lines).
__global__ void __launch_bounds__(ING_THREADS) k_ing_classify(IngText tx, const long long *starts, long long n_lines,
                                                             unsigned char *kind, int *text_of,
                                                             unsigned long long *first_endmdl,
                                                             unsigned long long *first_mmcif)
{
    const long long li = (long long)blockIdx.x * ING_THREADS + threadIdx.x;
    if (li >= n_lines) return;
    const long long s = starts[li], e = ing_line_end(tx, starts, n_lines, li);
    const unsigned char *p = tx.buf + s;
    const long long len = e - s;
    auto at = [&](int k) { return k < len ? p[k] : (unsigned char)0; };
    unsigned char k = ING_OTHER;
    if (at(0) == 'A' && at(1) == 'T' && at(2) == 'O' && at(3) == 'M') k = ING_ATOM;
    else if (at(0) == 'H' && at(1) == 'E' && at(2) == 'T' && at(3) == 'A' && at(4) == 'T' && at(5) == 'M') k = ING_HETATM;
    else if (at(0) == 'E' && at(1) == 'N' && at(2) == 'D' && at(3) == 'M' && at(4) == 'D' && at(5) == 'L') k = ING_ENDMDL;
    else if (len >= 5) {
        const unsigned char a = ing_lower(p[0]), b = ing_lower(p[1]), c = ing_lower(p[2]), d = ing_lower(p[3]);
        if (p[4] == '_' && ((a == 'd' && b == 'a' && c == 't' && d == 'a') || (a == 'l' && b == 'o' && c == 'o' && d == 'p')))
            k = ING_MMCIF;
    }
    kind[li] = k;
    const int t = ing_text_of(tx, s);
    text_of[li] = t;
    if (k == ING_ENDMDL) atomicMin(&first_endmdl[t], (unsigned long long)li);
    if (k == ING_MMCIF) atomicMin(&first_mmcif[t], (unsigned long long)li);
}

// whether line li is an atom record the reference would read
__device__ static inline int ing_is_atom(const IngText &tx, const unsigned char *kind, const int *text_of,
                                         const unsigned long long *first_endmdl, long long li)
{
    const unsigned char k = kind[li];
    if (k != ING_ATOM && !(k == ING_HETATM && !(tx.flags & 1))) return 0;
    // records after the first ENDMDL are not read unless it is ignored
    if (!(tx.flags & 2) && (unsigned long long)li > first_endmdl[text_of[li]]) return 0;
    return 1;
}

// atom records per block of ING_THREADS lines
__global__ void __launch_bounds__(ING_THREADS) k_ing_select_count(IngText tx, const unsigned char *kind,
                                                                 const int *text_of,
                                                                 const unsigned long long *first_endmdl,
                                                                 long long n_lines, long long *block_atoms)
{
    __shared__ long long lds[ING_THREADS / JAM_WAVE];
    const long long li = (long long)blockIdx.x * ING_THREADS + threadIdx.x;
    const long long v = li < n_lines ? ing_is_atom(tx, kind, text_of, first_endmdl, li) : 0;
    long long total;
    ing_block_scan(v, lds, total);
    if (threadIdx.x == 0) block_atoms[blockIdx.x] = total;
}

// atom_line[a]: the line of atom record a, in order; line_atom[li]: the atom
// index of line li's record, or of the next record (the count before it)
__global__ void __launch_bounds__(ING_THREADS) k_ing_select(IngText tx, const unsigned char *kind, const int *text_of,
                                                           const unsigned long long *first_endmdl, long long n_lines,
                                                           const long long *block_off, long long *atom_line,
                                                           long long *line_atom)
{
    __shared__ long long lds[ING_THREADS / JAM_WAVE];
    const long long li = (long long)blockIdx.x * ING_THREADS + threadIdx.x;
    const int v = li < n_lines ? ing_is_atom(tx, kind, text_of, first_endmdl, li) : 0;
    long long total;
    const long long a = block_off[blockIdx.x] + ing_block_scan(v, lds, total);
    if (li < n_lines) line_atom[li] = a;
    if (v) atom_line[a] = li;
}

// per text: its first line and its first atom (n_texts + 1 entries each)
__global__ void k_ing_text_ranges(IngText tx, const long long *starts, long long n_lines, const long long *line_atom,
                                  long long n_atoms, long long *text_line, long long *text_atom)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > tx.n_texts) return;
    long long lo = 0, hi = n_lines;   // first line starting at or after the text's first byte
    const long long pos = tx.text_off[t];
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (starts[mid] < pos) lo = mid + 1;
        else hi = mid;
    }
    text_line[t] = lo;
    text_atom[t] = lo < n_lines ? line_atom[lo] : n_atoms;
}

// ---------------------------------------------------------------------------
// fixed columns of one record
// ---------------------------------------------------------------------------

__constant__ double ING_POW10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                     1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

// str.strip()'s whitespace among ASCII bytes
__device__ static inline bool ing_space(unsigned char c) { return c == ' ' || (c >= 9 && c <= 13) || (c >= 28 && c <= 31); }

// int(field): optional sign, at least one digit, whitespace around
__device__ static inline bool ing_int(const unsigned char *f, int n, int &out)
{
    int i = 0;
    while (i < n && ing_space(f[i])) i++;
    bool neg = false;
    if (i < n && (f[i] == '+' || f[i] == '-')) neg = f[i++] == '-';
    long long v = 0;
    int digits = 0;
    while (i < n && f[i] >= '0' && f[i] <= '9') {
        v = v * 10 + (f[i++] - '0');
        if (++digits > 9) return false;   // fields are at most 5 characters
    }
    while (i < n && ing_space(f[i])) i++;
    if (digits == 0 || i != n) return false;
    out = (int)(neg ? -v : v);
    return true;
}

// float(field) for decimal forms: sign, digits with at most one point,
// optional exponent, whitespace around; exact (correctly rounded) when the
// significant digits fit 19 and the power of ten 22, else false
__device__ static inline bool ing_f64(const unsigned char *f, int n, double &out)
{
    int i = 0;
    while (i < n && ing_space(f[i])) i++;
    bool neg = false;
    if (i < n && (f[i] == '+' || f[i] == '-')) neg = f[i++] == '-';
    unsigned long long N = 0;
    int sig = 0, digits = 0, scale = 0;
    bool point = false;
    for (; i < n; i++) {
        const unsigned char c = f[i];
        if (c >= '0' && c <= '9') {
            digits++;
            if (N == 0 && c == '0') {
                if (point) scale--;
                continue;                      // leading zeros are not significant
            }
            if (++sig > 19) return false;
            N = N * 10 + (c - '0');
            if (point) scale--;
        } else if (c == '.' && !point) {
            point = true;
        } else {
            break;
        }
    }
    if (digits == 0) return false;
    if (i < n && (f[i] == 'e' || f[i] == 'E')) {
        i++;
        bool eneg = false;
        if (i < n && (f[i] == '+' || f[i] == '-')) eneg = f[i++] == '-';
        int ev = 0, ed = 0;
        while (i < n && f[i] >= '0' && f[i] <= '9') {
            ev = ev * 10 + (f[i++] - '0');
            if (++ed > 4) return false;
        }
        if (ed == 0) return false;
        scale += eneg ? -ev : ev;
    }
    while (i < n && ing_space(f[i])) i++;
    if (i != n) return false;
    double v;
    if (N == 0) {
        v = 0.0;
    } else {
        if (N >= (1ull << 53) || scale > 22 || scale < -22) return false;
        v = scale >= 0 ? (double)N * ING_POW10[scale] : (double)N / ING_POW10[-scale];
    }
    out = neg ? -v : v;
    return true;
}

// the per-atom arrays a parse fills
struct IngAtoms {
    double *xyz;             // 3 per atom
    uint32_t *name;          // 4 bytes, blanks and NUL as '_'
    uint32_t *resname;       // 3 bytes + NUL, blanks and NUL as '_'
    int *chain;              // chainID2 as a signed char value (what the search reads)
    unsigned char *chains;   // 2 per atom: chainID1, chainID2
    int *resseq, *serial;
    unsigned char *altloc, *icode;
    float *occupancy, *tempfactor;
    uint32_t *segid;         // 4 bytes, blanks and NUL as '_'
    unsigned short *element; // 2 bytes, blanks and NUL as '_'
    int *charge;
    unsigned int *err;       // per atom: ING_E* bits
    unsigned long long *first_err;   // device word: the first atom with an error (preset to n_atoms)
};

__device__ static inline unsigned char ing_tok(unsigned char c) { return (c == ' ' || c == 0) ? '_' : c; }

// Thread per atom record: Atom.loads (the PDB fixed columns, the line padded
// with blanks to 80 characters).
__global__ void __launch_bounds__(ING_THREADS) k_ing_parse(IngText tx, const long long *starts, long long n_lines,
                                                          const long long *atom_line, long long n_atoms, IngAtoms at)
{
    const long long a = (long long)blockIdx.x * ING_THREADS + threadIdx.x;
    if (a >= n_atoms) return;
    const long long li = atom_line[a];
    const long long s = starts[li], e = ing_line_end(tx, starts, n_lines, li);
    unsigned char col[80];
    unsigned int err = 0;
    const long long len = e - s;
#pragma unroll 8
    for (int k = 0; k < 80; k++) {
        const unsigned char c = k < len ? tx.buf[s + k] : (unsigned char)' ';
        if (c > 0x7F) err |= ING_ENONASCII;
        col[k] = c;
    }
    int serial = 0, resseq = 0;
    double x = 0.0, y = 0.0, z = 0.0, occ = 0.0, tf = 0.0;
    if (!ing_int(col + 6, 5, serial) || !ing_int(col + 22, 4, resseq) || !ing_f64(col + 30, 8, x) ||
        !ing_f64(col + 38, 8, y) || !ing_f64(col + 46, 8, z))
        err |= ING_EBADFIELD;
    // occupancy and temperature factor: blank is 0.0
    auto num = [&](int c0, int w, double &v) {
        int i = c0, j = c0 + w;
        while (i < j && ing_space(col[i])) i++;
        while (j > i && ing_space(col[j - 1])) j--;
        if (i == j) {
            v = 0.0;
            return true;
        }
        return ing_f64(col + i, j - i, v);
    };
    if (!num(54, 6, occ) || !num(60, 6, tf)) err |= ING_EBADFIELD;
    // charge: the stripped two columns read backwards ("2+" is +2), 0 when
    // that is not an integer
    int charge = 0;
    {
        int i = 78, j = 80;
        while (i < j && ing_space(col[i])) i++;
        while (j > i && ing_space(col[j - 1])) j--;
        unsigned char rev[2];
        const int m = j - i;
        for (int k = 0; k < m; k++) rev[k] = col[j - 1 - k];
        int v = 0;
        if (m > 0 && ing_int(rev, m, v)) charge = v;
    }
    at.xyz[3 * a + 0] = x;
    at.xyz[3 * a + 1] = y;
    at.xyz[3 * a + 2] = z;
    at.name[a] = ing_tok(col[12]) | (ing_tok(col[13]) << 8) | (ing_tok(col[14]) << 16) | ((uint32_t)ing_tok(col[15]) << 24);
    at.resname[a] = ing_tok(col[17]) | (ing_tok(col[18]) << 8) | (ing_tok(col[19]) << 16);
    at.chain[a] = (int)(signed char)col[21];
    at.chains[2 * a + 0] = col[20];
    at.chains[2 * a + 1] = col[21];
    at.resseq[a] = resseq;
    at.serial[a] = serial;
    at.altloc[a] = col[16];
    at.icode[a] = col[26];
    at.occupancy[a] = (float)occ;
    at.tempfactor[a] = (float)tf;
    at.segid[a] = ing_tok(col[72]) | (ing_tok(col[73]) << 8) | (ing_tok(col[74]) << 16) | ((uint32_t)ing_tok(col[75]) << 24);
    at.element[a] = (unsigned short)(ing_tok(col[76]) | (ing_tok(col[77]) << 8));
    at.charge[a] = charge;
    at.err[a] = err;
    if (err) atomicMin(at.first_err, (unsigned long long)a);
}

// ---------------------------------------------------------------------------
// residue-name index
// ---------------------------------------------------------------------------

// Block per molecule: its distinct upper-cased residue names, sorted, into
// keys[m * ING_MAX_NAMES ...] and their number into n_names[m] (more than
// ING_MAX_NAMES: n_names[m] = -1).
__global__ void __launch_bounds__(1024) k_ing_names(const int *atom_off, const uint32_t *resname, uint32_t *keys,
                                                    int *n_names)
{
    constexpr int SLOTS = 2 * ING_MAX_NAMES;   // open addressing at half load at most
    __shared__ uint32_t slot[SLOTS];
    __shared__ uint32_t list[ING_MAX_NAMES];
    __shared__ int count, overflow;
    const int m = blockIdx.x, t = threadIdx.x;
    for (int k = t; k < SLOTS; k += blockDim.x) slot[k] = 0xFFFFFFFFu;   // never an upper-cased name
    if (t == 0) {
        count = 0;
        overflow = 0;
    }
    __syncthreads();
    const int a0 = atom_off[m], a1 = atom_off[m + 1];
    for (int a = a0 + t; a < a1; a += blockDim.x) {
        const uint32_t key = jam_upper4(resname[a]);
        uint32_t h = (key * 2654435761u) & (SLOTS - 1);
        for (int probe = 0; probe < SLOTS; probe++) {
            const uint32_t old = atomicCAS(&slot[h], 0xFFFFFFFFu, key);
            if (old == 0xFFFFFFFFu) {
                const int c = atomicAdd(&count, 1);
                if (c < ING_MAX_NAMES) list[c] = key;
                else overflow = 1;
                break;
            }
            if (old == key) break;
            h = (h + 1) & (SLOTS - 1);
        }
    }
    __syncthreads();
    const int n = count;
    if (overflow || n > ING_MAX_NAMES) {
        if (t == 0) n_names[m] = -1;
        return;
    }
    // bitonic sort of the n names (padded to a power of two with all-ones)
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int k = n + t; k < n2; k += blockDim.x) list[k] = 0xFFFFFFFFu;
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < n2; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t u = list[i], v = list[ixj];
                    const bool up = (i & k) == 0;
                    if ((u > v) == up) {
                        list[i] = v;
                        list[ixj] = u;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int k = t; k < n; k += blockDim.x) keys[(size_t)m * ING_MAX_NAMES + k] = list[k];
    if (t == 0) n_names[m] = n;
}

// res_key[res_off[m] + k] = the k-th name of molecule m (res_off: scan of n_names)
__global__ void k_ing_keys(const uint32_t *keys, const int *n_names, const long long *res_off, int n_mols,
                           uint32_t *res_key)
{
    const int m = blockIdx.x;
    if (m >= n_mols) return;
    for (int k = threadIdx.x; k < n_names[m]; k += blockDim.x)
        res_key[res_off[m] + k] = keys[(size_t)m * ING_MAX_NAMES + k];
}

// bucket (global index) of every atom, and the bucket sizes: one atomic add
// per distinct bucket of a wave (atoms of a residue are adjacent, so a wave
// holds few distinct names) instead of one per atom on a few hot words
__global__ void __launch_bounds__(ING_THREADS) k_ing_bucket(const int *atom_off, int n_mols, long long n_atoms,
                                                           const uint32_t *resname, const uint32_t *res_key,
                                                           const long long *res_off, int *bucket, int *bucket_count)
{
    const long long a = (long long)blockIdx.x * ING_THREADS + threadIdx.x;
    const int lane = lane_id();
    long long b0 = 0;
    bool pending = a < n_atoms;
    if (pending) {
        int lo = 0, hi = n_mols;   // molecule of atom a: the last m with atom_off[m] <= a
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (atom_off[mid] <= a) lo = mid;
            else hi = mid;
        }
        const uint32_t key = jam_upper4(resname[a]);
        long long b1 = res_off[lo + 1];
        b0 = res_off[lo];
        while (b0 < b1) {
            const long long mid = (b0 + b1) >> 1;
            if (res_key[mid] < key) b0 = mid + 1;
            else b1 = mid;
        }
        bucket[a] = (int)b0;
    }
    // every lane takes part in the ballots and shuffles; each round retires
    // the lanes of the lowest pending lane's bucket (at most 64 rounds)
    for (int round = 0; round < JAM_WAVE; round++) {
        const unsigned long long m = ballot(pending);
        if (m == 0ull) break;
        const int leader = __ffsll((long long)m) - 1;
        const long long lb = __shfl(b0, leader, JAM_WAVE);
        const bool same = pending && b0 == lb;
        const unsigned long long peers = ballot(same);
        if (lane == leader) atomicAdd(&bucket_count[lb], __popcll(peers));
        if (same) pending = false;
    }
}

// Sort key of every atom: its bucket, then its serial number, which the
// stable radix sort orders as ResIndex does (upper-cased residue name, then
// serial, input order among equals); the value is the atom's index.
__global__ void __launch_bounds__(ING_THREADS) k_ing_sortkeys(long long n_atoms, const int *bucket, const int *serial,
                                                             unsigned long long *key, int *val)
{
    const long long a = (long long)blockIdx.x * ING_THREADS + threadIdx.x;
    if (a >= n_atoms) return;
    key[a] = ((unsigned long long)(unsigned int)bucket[a] << 32) | ((unsigned int)serial[a] ^ 0x80000000u);
    val[a] = (int)a;
}

// perm[i]: the molecule-local index of the i-th atom in sorted order (the
// buckets are molecule-major, so position i lies in its atom's molecule)
__global__ void __launch_bounds__(ING_THREADS) k_ing_perm(const int *atom_off, int n_mols, long long n_atoms,
                                                         const int *sorted, int *perm)
{
    const long long i = (long long)blockIdx.x * ING_THREADS + threadIdx.x;
    if (i >= n_atoms) return;
    const int a = sorted[i];
    int lo = 0, hi = n_mols;   // the last m with atom_off[m] <= a
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (atom_off[mid] <= a) lo = mid;
        else hi = mid;
    }
    perm[i] = a - atom_off[lo];
}

// res_begin / res_end of every bucket: its range in its molecule's part of
// perm (bucket_start: the exclusive scan of the bucket sizes, nb + 1 values)
__global__ void __launch_bounds__(ING_THREADS) k_ing_ranges(const int *atom_off, const long long *res_off, int n_mols,
                                                           long long n_buckets, const long long *bucket_start,
                                                           int *res_begin, int *res_end)
{
    const long long b = (long long)blockIdx.x * ING_THREADS + threadIdx.x;
    if (b >= n_buckets) return;
    int lo = 0, hi = n_mols;   // the last m with res_off[m] <= b: bucket b's molecule
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (res_off[mid] <= b) lo = mid;
        else hi = mid;
    }
    res_begin[b] = (int)(bucket_start[b] - atom_off[lo]);
    res_end[b] = (int)(bucket_start[b + 1] - atom_off[lo]);
}

}  // namespace jam
