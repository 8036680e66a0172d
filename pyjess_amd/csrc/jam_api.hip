// jam_api.hip -- host side of the C ABI declared in include/jess_amd.h.
//
// A query runs as one asynchronous chain on the context's stream:
//   k_candidates<count> -> exclusive scan -> k_candidates<fill> -> k_kd*
//   -> k_join -> k_super
// followed by one copy of the counters. Buffers are sized from the previous
// query; a kernel that would overflow one skips its writes (the counters keep
// counting) and the host grows the buffer to the exact size and re-runs the
// chain, so results never depend on capacities.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/jess_amd.h"
#include "jam_common.h"
#include "jam_grad.h"
#include "jam_ingest.h"
#include "jam_kernels.h"
#include "jam_sort.h"

namespace {

thread_local std::string g_error;

int32_t fail(int32_t code, const std::string &msg)
{
    g_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return fail(JAM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// A device allocation that grows (never shrinks) and frees itself.
struct DBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    hipError_t reserve(size_t n)
    {
        if (n <= bytes && ptr) return hipSuccess;
        release();
        size_t want = std::max<size_t>(n, 256);
        hipError_t e = hipMalloc(&ptr, want);
        if (e == hipSuccess) bytes = want;
        else ptr = nullptr;
        return e;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

template <class T>
hipError_t upload(DBuf &b, const std::vector<T> &v, hipStream_t s)
{
    hipError_t e = b.reserve(std::max<size_t>(1, v.size()) * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    e = hipMemcpyAsync(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

uint32_t pack4(const char *p)
{
    return (uint32_t)(uint8_t)p[0] | ((uint32_t)(uint8_t)p[1] << 8) | ((uint32_t)(uint8_t)p[2] << 16) |
           ((uint32_t)(uint8_t)p[3] << 24);
}

}  // namespace

// ---------------------------------------------------------------------------
// device scratch of the allocation-free entry point
// ---------------------------------------------------------------------------

// jam_superpose_batch_noalloc stages its inputs and outputs here instead of in
// buffers the library allocates: one copy of the array per device, held
// until the process ends. In host code `scratch` only names the symbol; its
// address on each device comes from hipGetSymbolAddress at context creation
// (g_scratch), and only that address is ever handed to the device.
static __device__ __align__(256) unsigned char scratch[JAM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int JAM_MAX_DEVICES = 64;
void *g_scratch[JAM_MAX_DEVICES] = {};
std::mutex g_scratch_init;                    // guards g_scratch
std::mutex g_scratch_busy[JAM_MAX_DEVICES];   // one call at a time uses a device's array
}  // namespace

// ---------------------------------------------------------------------------
// objects
// ---------------------------------------------------------------------------

struct jam_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    // start/stop events around every k_join launch (one per join level), for
    // the join kernels' own duration
    hipEvent_t xev[2 * JAM_MAX_LEVELS] = {};
    bool profile = false;       // join launches carry start/stop events (xev)
    // per-query workspace
    DBuf cnt, off, list, kidx, l2k, kent, katom, bbox, sbox, kscr, kd_bucket, kd_tied, kd_count;
    DBuf p_ord, p_lvl, p_nchunk, p_n1, p_base, p_cm, p_first, i_pair;
    // join: two frontiers (pair + stride entries per row), complete matches,
    // counters (rows per level, complete matches, overflow flag)
    DBuf j_pair[2], j_ent[2], d_pair, d_ent, j_count;
    int64_t front_cap = 0, done_cap = 0;
    int front_stride = 0, done_stride = 0;
    int join_grid[4] = {0, 0, 0, 0};   // resident workgroups of the join kernel <8/16/24/32>
    int kd_grid[14] = {};           // resident workgroups of the kd kernels
    // the exact kd pass of each size tier runs on a stream of its own, from
    // the moment the tier's fast pass has listed its tied sets, beside the
    // fast passes of the other tiers (all streams at the default priority)
    hipStream_t kd_side[KD_TIERS] = {};
    hipEvent_t kd_fast[KD_TIERS] = {}, kd_joined[KD_TIERS] = {};
    // the largest tier's sets of up to 8 * KD_LDS_SMALL_THREADS entries run
    // on a stream of their own, beside its larger sets
    hipStream_t kd_lds = nullptr;
    hipEvent_t kd_lds_done = nullptr;
    bool kd_pending = false;    // side-stream work of the last kd launch may still run
    // sort of the complete matches
    DBuf s_key[2], s_perm[2], s_pkey[2];
    DBuf s_starts;   // per template size: complete matches of at most that many atoms
    // radix-sort histograms and their scan; scan temporaries of the query
    DBuf rs_hist, rs_off;
    std::vector<std::unique_ptr<DBuf>> q_scan;
    int64_t sort_cap = 0;
    DBuf e_all;  // the events of the last query, event_layout(count, stride)
    DBuf keep_d, keep_i;  // superpositions kept between the two k_super passes (few matches)
    int64_t set_cap = 0, blk_cap = 0, sblk_cap = 0, event_cap = 0;
    int nsets_cap = 0, npairs_cap = 0;
    int64_t items_cap = 0;
    int event_stride = 0;
    // incremented by every query: a result's events live in e_* until the
    // next query on the context
    uint64_t query_gen = 0;
    // pinned host staging for event copies: two halves of JAM_STAGE_BYTES,
    // device-to-host copies into one half while the other is unpacked
    void *hstage = nullptr;
    hipEvent_t stage_ev[2] = {};
    // pinned host words the device's counters are read back into
    unsigned long long *hsmall = nullptr;
    int64_t hsmall_cap = 0;
    // standalone superposition buffers
    DBuf sa, sb, srmsd, srot, sc0, sc1;
    // standalone kd order: points, per-set offsets, zero chain/resSeq words
    DBuf o_xyz, o_aoff, o_zero;
    // per-pair hit counts and digests of jam_result_pair_digests
    DBuf dig;
    // temporary buffers of jam_molecules_parse_pdb, kept for the next parse
    DBuf g_buf, g_toff, g_nl, g_blk, g_blkoff, g_starts, g_kind, g_textof, g_first, g_lineatom, g_atomline, g_textrng;
    DBuf g_err, g_firsterr, g_keys, g_nnames, g_resoff64, g_bucket, g_bcount, g_bstart, g_skey, g_sval;
    std::vector<std::unique_ptr<DBuf>> g_scan;
};

struct jam_templates {
    jam_context *ctx = nullptr;
    int n_templates = 0, n_types = 0, stride = 1;
    DBuf atom_off, atom_type, chain, resseq, pos, weight, dist_off, dist;
    DBuf type_mode, type_name_off, type_names, type_rn_off, type_rns;
    DevTpls dev() const
    {
        DevTpls d;
        d.n_templates = n_templates;
        d.n_types = n_types;
        d.stride = stride;
        d.atom_off = atom_off.as<int>();
        d.atom_type = atom_type.as<int>();
        d.chain = chain.as<int>();
        d.resseq = resseq.as<int>();
        d.pos = pos.as<double>();
        d.weight = weight.as<double>();
        d.dist_off = dist_off.as<int>();
        d.dist = dist.as<double>();
        d.type_mode = type_mode.as<int>();
        d.type_name_off = type_name_off.as<int>();
        d.type_names = type_names.as<uint32_t>();
        d.type_rn_off = type_rn_off.as<int>();
        d.type_rns = type_rns.as<uint32_t>();
        return d;
    }
};

struct jam_molecules {
    jam_context *ctx = nullptr;
    int n_mols = 0;
    int64_t n_atoms = 0;
    int64_t n_buckets = 0;
    DBuf atom_off, xyz, name, resname, chain, resseq, perm, res_off, res_key, res_begin, res_end;
    // the other atom fields, kept by jam_molecules_parse_pdb for the host's copy
    bool parsed = false;
    DBuf serial, chains, altloc, icode, occupancy, tempfactor, segid, element, charge;
    DevMols dev() const
    {
        DevMols d;
        d.n_mols = n_mols;
        d.atom_off = atom_off.as<int>();
        d.xyz = xyz.as<double>();
        d.name = name.as<uint32_t>();
        d.resname = resname.as<uint32_t>();
        d.chain = chain.as<int>();
        d.resseq = resseq.as<int>();
        d.perm = perm.as<int>();
        d.res_off = res_off.as<int>();
        d.res_key = res_key.as<uint32_t>();
        d.res_begin = res_begin.as<int>();
        d.res_end = res_end.as<int>();
        return d;
    }
};

struct jam_result {
    int n_pairs = 0;
    int stride = 1;
    // the events stay in the context's device buffers (copied out on demand)
    jam_context *ctx = nullptr;
    uint64_t gen = 0;
    int64_t n_events = 0;
    std::vector<int64_t> pair_cm;
    double ms[5] = {0, 0, 0, 0, 0};
    double expand_ms = 0.0;     // summed duration of the k_join launches
    int64_t level_partials[JAM_MAX_LEVELS] = {};   // partial matches expanded per join level
    int expand_launches = 0;
};

// ---------------------------------------------------------------------------
// superposition RMSD and its gradient on device arrays
// ---------------------------------------------------------------------------

extern "C" int32_t jam_rmsd_forward(const double *a, const double *b, const int64_t *off, int64_t n_sets,
                                    int64_t n_points, double *rmsd, double *rot9, void *stream)
{
    if (n_sets < 0 || n_points < 0) return fail(JAM_EINVAL, "invalid sizes");
    if (n_sets == 0) return JAM_OK;
    if (!off || !rmsd || (n_points > 0 && (!a || !b))) return fail(JAM_EINVAL, "NULL argument");
    const int tb = RG_SETS;
    hipLaunchKernelGGL(jam::k_rmsd_fwd, dim3((unsigned)((n_sets + tb - 1) / tb)), dim3(tb), 0,
                       (hipStream_t)stream, a, b, off, (long long)n_sets, (long long)n_points, rmsd, rot9);
    HIP_TRY(hipGetLastError());
    return JAM_OK;
}

extern "C" int32_t jam_rmsd_backward(const double *a, const double *b, const int64_t *off, int64_t n_sets,
                                     int64_t n_points, const double *rmsd, const double *rot9,
                                     const double *grad_rmsd, double *grad_a, double *grad_b, void *stream)
{
    if (n_sets < 0 || n_points < 0) return fail(JAM_EINVAL, "invalid sizes");
    if (n_sets == 0) return JAM_OK;
    if (!off || !rmsd || !grad_rmsd || (!grad_a && !grad_b) || (n_points > 0 && (!a || !b)))
        return fail(JAM_EINVAL, "NULL argument");
    const int tb = RG_SETS;
    hipLaunchKernelGGL(jam::k_rmsd_bwd, dim3((unsigned)((n_sets + tb - 1) / tb)), dim3(tb), 0,
                       (hipStream_t)stream, a, b, off, (long long)n_sets, (long long)n_points, rmsd, rot9,
                       grad_rmsd, grad_a, grad_b);
    HIP_TRY(hipGetLastError());
    return JAM_OK;
}

// ---------------------------------------------------------------------------
// misc
// ---------------------------------------------------------------------------

extern "C" const char *jam_last_error(void) { return g_error.c_str(); }

extern "C" const char *jam_version(void) { return "pyjess-amd 0.1.0 (gfx950)"; }

static bool is_gfx950(int dev)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return false;
    return std::strncmp(p.gcnArchName, "gfx950", 6) == 0;
}

extern "C" int32_t jam_device_available(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    for (int i = 0; i < n; i++)
        if (is_gfx950(i)) return 1;
    return 0;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

extern "C" int32_t jam_context_create(int32_t device, jam_context **out)
{
    if (!out) return fail(JAM_EINVAL, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(JAM_EDEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(JAM_EINVAL, "device ordinal out of range");
    if (!is_gfx950(device)) return fail(JAM_EDEVICE, "device is not a gfx950 (MI355X) GPU");
    HIP_TRY(hipSetDevice(device));
    if (device < JAM_MAX_DEVICES) {
        // the device address of this device's scratch array, once per device
        std::lock_guard<std::mutex> g(g_scratch_init);
        if (!g_scratch[device]) {
            void *p = nullptr;
            HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scratch[device] = p;
        }
    }
    std::unique_ptr<jam_context> c(new jam_context());
    c->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &ev : c->ev) HIP_TRY(hipEventCreate(&ev));
    for (auto &ev : c->xev) HIP_TRY(hipEventCreate(&ev));
    for (int k = 0; k < KD_TIERS; k++) {
        HIP_TRY(hipStreamCreateWithFlags(&c->kd_side[k], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&c->kd_fast[k], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&c->kd_joined[k], hipEventDisableTiming));
    }
    HIP_TRY(hipStreamCreateWithFlags(&c->kd_lds, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->kd_lds_done, hipEventDisableTiming));
    *out = c.release();
    return JAM_OK;
}

extern "C" int32_t jam_synchronize(jam_context *ctx)
{
    if (!ctx) return fail(JAM_EINVAL, "NULL context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    return JAM_OK;
}

extern "C" void jam_context_free(jam_context *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto &ev : ctx->xev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->hstage) (void)hipHostFree(ctx->hstage);
    if (ctx->hsmall) (void)hipHostFree(ctx->hsmall);
    for (auto &ev : ctx->stage_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (int k = 0; k < KD_TIERS; k++) {
        if (ctx->kd_side[k]) (void)hipStreamSynchronize(ctx->kd_side[k]);
        if (ctx->kd_fast[k]) (void)hipEventDestroy(ctx->kd_fast[k]);
        if (ctx->kd_joined[k]) (void)hipEventDestroy(ctx->kd_joined[k]);
        if (ctx->kd_side[k]) (void)hipStreamDestroy(ctx->kd_side[k]);
    }
    if (ctx->kd_lds) (void)hipStreamSynchronize(ctx->kd_lds);
    if (ctx->kd_lds_done) (void)hipEventDestroy(ctx->kd_lds_done);
    if (ctx->kd_lds) (void)hipStreamDestroy(ctx->kd_lds);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

// ---------------------------------------------------------------------------
// templates
// ---------------------------------------------------------------------------

extern "C" int32_t jam_templates_create(jam_context *ctx, const jam_template_desc *d, jam_templates **out)
{
    if (!ctx || !d || !out) return fail(JAM_EINVAL, "NULL argument");
    *out = nullptr;
    if (d->n_templates < 0) return fail(JAM_EINVAL, "negative template count");
    const int T = d->n_templates;
    const int A = T > 0 ? d->atom_off[T] : 0;
    for (int t = 0; t < T; t++) {
        int n = d->atom_off[t + 1] - d->atom_off[t];
        if (n < 0) return fail(JAM_EINVAL, "atom offsets are not increasing");
        if (n > JAM_MAX_TEMPLATE_ATOMS)
            return fail(JAM_ETOOBIG, "template " + std::to_string(t) + " has " + std::to_string(n) +
                                         " atoms, more than JAM_MAX_TEMPLATE_ATOMS");
    }
    for (int a = 0; a < A; a++) {
        int m = d->mode[a];
        if (m != 0 && m != 1 && m != 3)
            return fail(JAM_EMODE, "match mode " + std::to_string(m) +
                                       " is not implemented on the device (implemented: 0, 1, 3)");
    }
    // atom types: (mode, names, residue names in order), deduplicated
    std::map<std::tuple<int, std::vector<uint32_t>, std::vector<uint32_t>>, int> types;
    std::vector<int> atom_type(A), type_mode, type_name_off{0}, type_rn_off{0};
    std::vector<uint32_t> type_names, type_rns;
    for (int a = 0; a < A; a++) {
        std::vector<uint32_t> nm, rn;
        for (int k = d->name_off[a]; k < d->name_off[a + 1]; k++) nm.push_back(pack4(d->names + 4 * k));
        for (int k = d->resname_off[a]; k < d->resname_off[a + 1]; k++)
            rn.push_back(jam_upper4(pack4(d->resnames + 4 * k)));
        auto key = std::make_tuple(d->mode[a], nm, rn);
        auto it = types.find(key);
        if (it == types.end()) {
            int id = (int)type_mode.size();
            types.emplace(key, id);
            type_mode.push_back(d->mode[a]);
            type_names.insert(type_names.end(), nm.begin(), nm.end());
            type_name_off.push_back((int)type_names.size());
            type_rns.insert(type_rns.end(), rn.begin(), rn.end());
            type_rn_off.push_back((int)type_rns.size());
            atom_type[a] = id;
        } else {
            atom_type[a] = it->second;
        }
    }
    std::vector<int> atom_off(d->atom_off, d->atom_off + T + 1), chain(A), resseq(A), dist_off(T);
    std::vector<double> pos(d->pos, d->pos + 3 * (size_t)A), weight(d->weight, d->weight + A), dist;
    int stride = 1;
    for (int a = 0; a < A; a++) {
        chain[a] = (int)(signed char)d->chain2[a];
        resseq[a] = d->resseq[a];
    }
    // template distances: TessAtom_distance, sqrt of squared differences summed x, y, z
    for (int t = 0; t < T; t++) {
        int b = atom_off[t], n = atom_off[t + 1] - b;
        stride = std::max(stride, n);
        dist_off[t] = (int)dist.size();
        for (int i = 0; i < n; i++)
            for (int k = 0; k < n; k++) {
                double s = 0.0;
                for (int q = 0; q < 3; q++) {
                    double v = pos[3 * (b + i) + q] - pos[3 * (b + k) + q];
                    s += v * v;
                }
                dist.push_back(i == k ? 0.0 : std::sqrt(s));
            }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<jam_templates> t(new jam_templates());
    t->ctx = ctx;
    t->n_templates = T;
    t->n_types = (int)type_mode.size();
    t->stride = stride;
    hipStream_t s = ctx->stream;
    HIP_TRY(upload(t->atom_off, atom_off, s));
    HIP_TRY(upload(t->atom_type, atom_type, s));
    HIP_TRY(upload(t->chain, chain, s));
    HIP_TRY(upload(t->resseq, resseq, s));
    HIP_TRY(upload(t->pos, pos, s));
    HIP_TRY(upload(t->weight, weight, s));
    HIP_TRY(upload(t->dist_off, dist_off, s));
    HIP_TRY(upload(t->dist, dist, s));
    HIP_TRY(upload(t->type_mode, type_mode, s));
    HIP_TRY(upload(t->type_name_off, type_name_off, s));
    HIP_TRY(upload(t->type_names, type_names, s));
    HIP_TRY(upload(t->type_rn_off, type_rn_off, s));
    HIP_TRY(upload(t->type_rns, type_rns, s));
    *out = t.release();
    return JAM_OK;
}

extern "C" void jam_templates_free(jam_templates *t)
{
    if (!t) return;
    (void)hipSetDevice(t->ctx->device);
    delete t;
}

// ---------------------------------------------------------------------------
// molecules
// ---------------------------------------------------------------------------

extern "C" int32_t jam_molecules_create(jam_context *ctx, const jam_molecule_desc *mols, int32_t n,
                                        jam_molecules **out)
{
    if (!ctx || !out || (n > 0 && !mols)) return fail(JAM_EINVAL, "NULL argument");
    if (n < 0) return fail(JAM_EINVAL, "negative molecule count");
    *out = nullptr;
    std::vector<int> atom_off{0}, chain, resseq, perm, res_off{0}, res_begin, res_end;
    std::vector<double> xyz;
    std::vector<uint32_t> name, resname, res_key;
    for (int m = 0; m < n; m++) {
        const jam_molecule_desc &M = mols[m];
        if (M.n_atoms < 0) return fail(JAM_EINVAL, "negative atom count");
        const int N = M.n_atoms;
        if ((int64_t)atom_off.back() + N > INT32_MAX) return fail(JAM_ETOOBIG, "too many atoms in one batch");
        xyz.insert(xyz.end(), M.xyz, M.xyz + 3 * (size_t)N);
        std::vector<uint32_t> upkey(N);
        for (int i = 0; i < N; i++) {
            name.push_back(pack4(M.names + 4 * i));
            uint32_t rn = pack4(M.resnames + 4 * i);
            resname.push_back(rn);
            upkey[i] = jam_upper4(rn);
            chain.push_back((int)(signed char)M.chain2[i]);
            resseq.push_back(M.resseq[i]);
        }
        // residue-name index (ResIndex_create, patches/ResIndex.c.patch:34-119):
        // atoms grouped by case-insensitive residue name, each group by serial,
        // ties in input order. Groups are ordered by their packed key value,
        // the order the device binary search expects; the order of groups has
        // no effect on candidate lists, only the order inside a group does.
        std::vector<int> p(N);
        std::iota(p.begin(), p.end(), 0);
        std::stable_sort(p.begin(), p.end(), [&](int a, int b) {
            if (upkey[a] != upkey[b]) return upkey[a] < upkey[b];
            return M.serial[a] < M.serial[b];
        });
        perm.insert(perm.end(), p.begin(), p.end());
        for (int i = 0; i < N; i++) {
            if (i == 0 || upkey[p[i]] != upkey[p[i - 1]]) {
                if (i > 0) res_end.push_back(i);
                res_key.push_back(upkey[p[i]]);
                res_begin.push_back(i);
            }
        }
        if (N > 0) res_end.push_back(N);
        res_off.push_back((int)res_key.size());
        atom_off.push_back(atom_off.back() + N);
    }
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<jam_molecules> M(new jam_molecules());
    M->ctx = ctx;
    M->n_mols = n;
    M->n_atoms = atom_off.back();
    hipStream_t s = ctx->stream;
    HIP_TRY(upload(M->atom_off, atom_off, s));
    HIP_TRY(upload(M->xyz, xyz, s));
    HIP_TRY(upload(M->name, name, s));
    HIP_TRY(upload(M->resname, resname, s));
    HIP_TRY(upload(M->chain, chain, s));
    HIP_TRY(upload(M->resseq, resseq, s));
    HIP_TRY(upload(M->perm, perm, s));
    HIP_TRY(upload(M->res_off, res_off, s));
    HIP_TRY(upload(M->res_key, res_key, s));
    HIP_TRY(upload(M->res_begin, res_begin, s));
    HIP_TRY(upload(M->res_end, res_end, s));
    M->n_buckets = (int64_t)res_key.size();
    *out = M.release();
    return JAM_OK;
}

extern "C" void jam_molecules_free(jam_molecules *m)
{
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    delete m;
}

// ---------------------------------------------------------------------------
// PDB ingestion on the device
// ---------------------------------------------------------------------------

namespace {

constexpr int SCAN_BLOCK = 1024;
// arrays of at most this many SCAN_BLOCK-element blocks are scanned by one
// workgroup in one launch
constexpr long long SCAN_ONE_BLOCKS = 32;

// Exclusive scan of 1024-element blocks of `in` into `out`, each block's
// total into `sums` (sums in 64 bits).
template <class T, class O>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_blocks(const T *in, long long n, O *out, long long *sums)
{
    __shared__ long long w[SCAN_BLOCK / JAM_WAVE];
    const long long i = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    const long long v = i < n ? (long long)in[i] : 0;
    const int lane = jam::lane_id(), wv = threadIdx.x / JAM_WAVE;
    long long x = v;
    for (int sh = 1; sh < JAM_WAVE; sh <<= 1) {
        const long long y = __shfl_up(x, sh, JAM_WAVE);
        if (lane >= sh) x += y;
    }
    if (lane == JAM_WAVE - 1) w[wv] = x;
    __syncthreads();
    long long base = 0, total = 0;
    for (int k = 0; k < SCAN_BLOCK / JAM_WAVE; k++) {
        if (k < wv) base += w[k];
        total += w[k];
    }
    if (i < n) out[i] = (O)(base + x - v);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// The whole exclusive scan in one workgroup, SCAN_BLOCK elements at a time
// with a running carry, the total into out[n]: for small arrays, where the
// launches of the multi-block scan would dominate.
template <class T, class O>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_one(const T *in, long long n, O *out)
{
    __shared__ long long w[SCAN_BLOCK / JAM_WAVE];
    const int lane = jam::lane_id(), wv = threadIdx.x / JAM_WAVE;
    long long carry = 0;
    for (long long b0 = 0; b0 < n; b0 += SCAN_BLOCK) {
        const long long i = b0 + threadIdx.x;
        const long long v = i < n ? (long long)in[i] : 0;
        long long x = v;
        for (int sh = 1; sh < JAM_WAVE; sh <<= 1) {
            const long long y = __shfl_up(x, sh, JAM_WAVE);
            if (lane >= sh) x += y;
        }
        if (lane == JAM_WAVE - 1) w[wv] = x;
        __syncthreads();
        long long base = 0, total = 0;
        for (int k = 0; k < SCAN_BLOCK / JAM_WAVE; k++) {
            if (k < wv) base += w[k];
            total += w[k];
        }
        if (i < n) out[i] = (O)(carry + base + x - v);
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = (O)carry;
}

template <class O>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_add(O *out, long long n, const long long *offs)
{
    const long long i = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    if (i < n) out[i] += (O)offs[blockIdx.x];
}

template <class O>
__global__ void k_scan_total(O *dst, const long long *src)
{
    if (threadIdx.x == 0) *dst = (O)*src;
}

__global__ void k_fill_u64(unsigned long long *v, long long n, unsigned long long x)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        v[i] = x;
}

// a line feed at each of the n positions of buf
__global__ void k_put_newlines(const long long *pos, long long n, unsigned char *buf)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) buf[pos[i]] = '\n';
}

__global__ void k_fill_i32(int *v, long long n, int x)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        v[i] = x;
}

// *out = the largest of in[0, n) (0 when n = 0): one workgroup
__global__ void __launch_bounds__(1024) k_max_i32(const int *in, long long n, int *out)
{
    __shared__ int w[1024 / JAM_WAVE];
    int m = 0;
    for (long long i = threadIdx.x; i < n; i += 1024) m = max(m, in[i]);
    for (int sh = JAM_WAVE / 2; sh >= 1; sh >>= 1) m = max(m, __shfl_xor(m, sh, JAM_WAVE));
    if (jam::lane_id() == 0) w[threadIdx.x / JAM_WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 1024 / JAM_WAVE; k++) m = max(m, w[k]);
        *out = max(m, w[0]);
    }
}

// out[0, n] = exclusive scan of in[0, n) followed by the total (n + 1
// values), on stream s; temporary space from `tmp`, grown as needed
template <class T, class O>
int32_t device_scan(const T *in, long long n, O *out, std::vector<std::unique_ptr<DBuf>> &tmp, size_t level,
                    hipStream_t s)
{
    const long long blocks = std::max<long long>(1, (n + SCAN_BLOCK - 1) / SCAN_BLOCK);
    if (blocks <= SCAN_ONE_BLOCKS) {
        hipLaunchKernelGGL((k_scan_one<T, O>), dim3(1), dim3(SCAN_BLOCK), 0, s, in, n, out);
        HIP_TRY(hipGetLastError());
        return JAM_OK;
    }
    while (tmp.size() <= level + 1) tmp.emplace_back(new DBuf());
    HIP_TRY(tmp[level]->reserve(sizeof(long long) * (size_t)(blocks + 1)));
    long long *sums = tmp[level]->as<long long>();
    hipLaunchKernelGGL((k_scan_blocks<T, O>), dim3((unsigned)blocks), dim3(SCAN_BLOCK), 0, s, in, n, out, sums);
    HIP_TRY(hipGetLastError());
    if (blocks == 1) {
        // the total goes after the scan
        hipLaunchKernelGGL(k_scan_total<O>, dim3(1), dim3(64), 0, s, out + n, (const long long *)sums);
        HIP_TRY(hipGetLastError());
        return JAM_OK;
    }
    // scan of the block totals (with their total at [blocks]), added back
    HIP_TRY(tmp[level + 1]->reserve(sizeof(long long) * (size_t)(blocks + 1)));
    long long *offs = tmp[level + 1]->as<long long>();
    int32_t rc = device_scan<long long, long long>(sums, blocks, offs, tmp, level + 2, s);
    if (rc != JAM_OK) return rc;
    hipLaunchKernelGGL(k_scan_add<O>, dim3((unsigned)blocks), dim3(SCAN_BLOCK), 0, s, out, n,
                       (const long long *)offs);
    hipLaunchKernelGGL(k_scan_total<O>, dim3(1), dim3(64), 0, s, out + n, (const long long *)(offs + blocks));
    HIP_TRY(hipGetLastError());
    return JAM_OK;
}

// Stable LSD radix sort of n (key, value) pairs by bits [0, bits) of the
// keys, 8 bits per pass, between the buffer pairs (k0, v0) -- the input --
// and (k1, v1); `which` is set to the pair holding the result (0 or 1).
template <class K>
int32_t radix_sort(K *k0, int *v0, K *k1, int *v1, long long n, int bits, DBuf &hist, DBuf &offs,
                   std::vector<std::unique_ptr<DBuf>> &tmp, hipStream_t s, int &which)
{
    which = 0;
    if (n <= 1 || bits <= 0) return JAM_OK;
    if (n >= INT32_MAX) return fail(JAM_ETOOBIG, "more than 2^31 keys in one sort");
    const long long tiles = (n + RS_TILE - 1) / RS_TILE;
    HIP_TRY(hist.reserve(sizeof(int) * (size_t)(RS_THREADS * tiles)));
    HIP_TRY(offs.reserve(sizeof(int) * (size_t)(RS_THREADS * tiles + 1)));
    K *ki = k0, *ko = k1;
    int *vi = v0, *vo = v1;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(jam::k_rs_hist<K>, dim3((unsigned)tiles), dim3(RS_THREADS), 0, s, (const K *)ki, n, shift,
                           hist.as<int>(), tiles);
        int32_t rc = device_scan<int, int>(hist.as<const int>(), RS_THREADS * tiles, offs.as<int>(), tmp, 0, s);
        if (rc != JAM_OK) return rc;
        hipLaunchKernelGGL(jam::k_rs_scatter<K>, dim3((unsigned)tiles), dim3(RS_THREADS), 0, s, (const K *)ki,
                           (const int *)vi, n, shift, offs.as<const int>(), tiles, ko, vo);
        HIP_TRY(hipGetLastError());
        std::swap(ki, ko);
        std::swap(vi, vo);
        which ^= 1;
    }
    return JAM_OK;
}

template <class T>
int32_t read_one(const T *dev, T &host, hipStream_t s)
{
    HIP_TRY(hipMemcpyAsync(&host, dev, sizeof(T), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return JAM_OK;
}

}  // namespace

extern "C" int32_t jam_molecules_parse_pdb(jam_context *c, const char *const *texts, const int64_t *lengths,
                                           int32_t n, int32_t flags, jam_molecules **out)
{
    if (!c || !out || n < 0 || (n > 0 && (!texts || !lengths))) return fail(JAM_EINVAL, "invalid argument");
    *out = nullptr;
    if (flags & ~(JAM_PDB_SKIP_HETATM | JAM_PDB_IGNORE_ENDMDL)) return fail(JAM_EINVAL, "unknown flags");
    // where the texts go on the device: each one's bytes, then a line feed
    // when it does not end with one (its last line ends there)
    std::vector<long long> toff(n + 1, 0), nl;
    for (int32_t k = 0; k < n; k++) {
        if (lengths[k] < 0 || (lengths[k] > 0 && !texts[k]))
            return fail(JAM_EINVAL, "text " + std::to_string(k) + ": negative length or NULL text");
        toff[k + 1] = toff[k] + lengths[k];
        if (lengths[k] > 0 && texts[k][lengths[k] - 1] != '\n') nl.push_back(toff[k + 1]++);
    }
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    std::unique_ptr<jam_molecules> M(new jam_molecules());
    M->ctx = c;
    M->n_mols = n;
    M->parsed = true;
    std::vector<std::unique_ptr<DBuf>> &tmp = c->g_scan;
    HIP_TRY(c->g_buf.reserve((size_t)std::max<long long>(toff[n], 1)));
    for (int32_t k = 0; k < n; k++)
        if (lengths[k] > 0)
            HIP_TRY(hipMemcpyAsync(c->g_buf.as<unsigned char>() + toff[k], texts[k], (size_t)lengths[k],
                                   hipMemcpyHostToDevice, s));
    if (!nl.empty()) {
        HIP_TRY(upload(c->g_nl, nl, s));
        hipLaunchKernelGGL(k_put_newlines, dim3((unsigned)((nl.size() + 255) / 256)), dim3(256), 0, s,
                           c->g_nl.as<const long long>(), (long long)nl.size(), c->g_buf.as<unsigned char>());
    }
    HIP_TRY(upload(c->g_toff, toff, s));
    jam::IngText tx;
    tx.buf = c->g_buf.as<const unsigned char>();
    tx.n_bytes = toff[n];
    tx.text_off = c->g_toff.as<const long long>();
    tx.n_texts = n;
    tx.flags = flags;
    const long long per_block = (long long)ING_THREADS * ING_BPT;
    const long long byte_blocks = std::max<long long>(1, (tx.n_bytes + per_block - 1) / per_block);
    // ---- lines
    HIP_TRY(c->g_blk.reserve(sizeof(long long) * (size_t)byte_blocks));
    HIP_TRY(c->g_blkoff.reserve(sizeof(long long) * (size_t)(byte_blocks + 1)));
    hipLaunchKernelGGL(jam::k_ing_count, dim3((unsigned)byte_blocks), dim3(ING_THREADS), 0, s, tx,
                       c->g_blk.as<long long>());
    int32_t rc;
    if ((rc = device_scan<long long>(c->g_blk.as<long long>(), byte_blocks, c->g_blkoff.as<long long>(), tmp, 0, s)) != JAM_OK)
        return rc;
    long long n_lines = 0;
    if ((rc = read_one(c->g_blkoff.as<long long>() + byte_blocks, n_lines, s)) != JAM_OK) return rc;
    if (tx.n_bytes == 0) n_lines = 0;
    const long long lines_cap = std::max<long long>(n_lines, 1);
    HIP_TRY(c->g_starts.reserve(sizeof(long long) * (size_t)lines_cap));
    if (n_lines > 0)
        hipLaunchKernelGGL(jam::k_ing_starts, dim3((unsigned)byte_blocks), dim3(ING_THREADS), 0, s, tx,
                           c->g_blkoff.as<const long long>(), c->g_starts.as<long long>());
    // ---- record kinds, first ENDMDL / mmCIF tag per text
    HIP_TRY(c->g_kind.reserve((size_t)lines_cap));
    HIP_TRY(c->g_textof.reserve(sizeof(int) * (size_t)lines_cap));
    HIP_TRY(c->g_first.reserve(sizeof(unsigned long long) * 2 * (size_t)std::max(n, 1)));
    unsigned long long *first_endmdl = c->g_first.as<unsigned long long>();
    unsigned long long *first_mmcif = first_endmdl + std::max(n, 1);
    hipLaunchKernelGGL(k_fill_u64, dim3(16), dim3(256), 0, s, first_endmdl, 2 * (long long)std::max(n, 1),
                       (unsigned long long)n_lines);
    const long long line_blocks = std::max<long long>(1, (n_lines + ING_THREADS - 1) / ING_THREADS);
    if (n_lines > 0)
        hipLaunchKernelGGL(jam::k_ing_classify, dim3((unsigned)line_blocks), dim3(ING_THREADS), 0, s, tx,
                           c->g_starts.as<const long long>(), n_lines, c->g_kind.as<unsigned char>(), c->g_textof.as<int>(),
                           first_endmdl, first_mmcif);
    // ---- atom records
    HIP_TRY(c->g_blk.reserve(sizeof(long long) * (size_t)line_blocks));
    HIP_TRY(c->g_blkoff.reserve(sizeof(long long) * (size_t)(line_blocks + 1)));
    HIP_TRY(c->g_lineatom.reserve(sizeof(long long) * (size_t)lines_cap));
    long long n_atoms = 0;
    if (n_lines > 0) {
        hipLaunchKernelGGL(jam::k_ing_select_count, dim3((unsigned)line_blocks), dim3(ING_THREADS), 0, s, tx,
                           c->g_kind.as<const unsigned char>(), c->g_textof.as<const int>(),
                           (const unsigned long long *)first_endmdl, n_lines, c->g_blk.as<long long>());
        if ((rc = device_scan<long long>(c->g_blk.as<long long>(), line_blocks, c->g_blkoff.as<long long>(), tmp, 0, s)) !=
            JAM_OK)
            return rc;
        if ((rc = read_one(c->g_blkoff.as<long long>() + line_blocks, n_atoms, s)) != JAM_OK) return rc;
    }
    if (n_atoms >= INT32_MAX) return fail(JAM_ETOOBIG, "too many atoms in one batch");
    const long long atoms_cap = std::max<long long>(n_atoms, 1);
    HIP_TRY(c->g_atomline.reserve(sizeof(long long) * (size_t)atoms_cap));
    if (n_lines > 0)
        hipLaunchKernelGGL(jam::k_ing_select, dim3((unsigned)line_blocks), dim3(ING_THREADS), 0, s, tx,
                           c->g_kind.as<const unsigned char>(), c->g_textof.as<const int>(),
                           (const unsigned long long *)first_endmdl, n_lines, c->g_blkoff.as<const long long>(),
                           c->g_atomline.as<long long>(), c->g_lineatom.as<long long>());
    HIP_TRY(c->g_textrng.reserve(sizeof(long long) * 2 * (size_t)(n + 1)));
    long long *text_line = c->g_textrng.as<long long>(), *text_atom = text_line + (n + 1);
    hipLaunchKernelGGL(jam::k_ing_text_ranges, dim3((unsigned)((n + 1 + 63) / 64)), dim3(64), 0, s, tx,
                       c->g_starts.as<const long long>(), n_lines, c->g_lineatom.as<const long long>(), n_atoms, text_line,
                       text_atom);
    HIP_TRY(hipGetLastError());
    std::vector<long long> h_rng(2 * (size_t)(n + 1)), h_first(2 * (size_t)std::max(n, 1));
    HIP_TRY(hipMemcpyAsync(h_rng.data(), text_line, sizeof(long long) * h_rng.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_first.data(), first_endmdl, sizeof(long long) * h_first.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // an mmCIF tag before the parse stops (PDBMoleculeParser raises)
    for (int32_t k = 0; k < n; k++) {
        const long long tag = h_first[std::max(n, 1) + k], end = h_first[k];
        if (tag < n_lines && ((flags & JAM_PDB_IGNORE_ENDMDL) || tag < end))
            return fail(JAM_EINVAL, "text " + std::to_string(k) + ": mmCIF data tags found, file is not in PDB format");
    }
    std::vector<int> aoff(n + 1);
    for (int32_t k = 0; k <= n; k++) aoff[k] = (int)h_rng[(size_t)(n + 1) + k];
    M->n_atoms = n_atoms;
    // ---- fields
    const size_t na = (size_t)atoms_cap;
    HIP_TRY(upload(M->atom_off, aoff, s));
    HIP_TRY(M->xyz.reserve(sizeof(double) * 3 * na));
    HIP_TRY(M->name.reserve(4 * na));
    HIP_TRY(M->resname.reserve(4 * na));
    HIP_TRY(M->chain.reserve(sizeof(int) * na));
    HIP_TRY(M->chains.reserve(2 * na));
    HIP_TRY(M->resseq.reserve(sizeof(int) * na));
    HIP_TRY(M->serial.reserve(sizeof(int) * na));
    HIP_TRY(M->altloc.reserve(na));
    HIP_TRY(M->icode.reserve(na));
    HIP_TRY(M->occupancy.reserve(sizeof(float) * na));
    HIP_TRY(M->tempfactor.reserve(sizeof(float) * na));
    HIP_TRY(M->segid.reserve(4 * na));
    HIP_TRY(M->element.reserve(2 * na));
    HIP_TRY(M->charge.reserve(sizeof(int) * na));
    HIP_TRY(c->g_err.reserve(sizeof(unsigned int) * na));
    HIP_TRY(c->g_firsterr.reserve(sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, c->g_firsterr.as<unsigned long long>(), 1ll,
                       (unsigned long long)n_atoms);
    jam::IngAtoms at;
    at.xyz = M->xyz.as<double>();
    at.name = M->name.as<uint32_t>();
    at.resname = M->resname.as<uint32_t>();
    at.chain = M->chain.as<int>();
    at.chains = M->chains.as<unsigned char>();
    at.resseq = M->resseq.as<int>();
    at.serial = M->serial.as<int>();
    at.altloc = M->altloc.as<unsigned char>();
    at.icode = M->icode.as<unsigned char>();
    at.occupancy = M->occupancy.as<float>();
    at.tempfactor = M->tempfactor.as<float>();
    at.segid = M->segid.as<uint32_t>();
    at.element = M->element.as<unsigned short>();
    at.charge = M->charge.as<int>();
    at.err = c->g_err.as<unsigned int>();
    at.first_err = c->g_firsterr.as<unsigned long long>();
    if (n_atoms > 0)
        hipLaunchKernelGGL(jam::k_ing_parse, dim3((unsigned)((n_atoms + ING_THREADS - 1) / ING_THREADS)),
                           dim3(ING_THREADS), 0, s, tx, c->g_starts.as<const long long>(), n_lines,
                           c->g_atomline.as<const long long>(), n_atoms, at);
    HIP_TRY(hipGetLastError());
    unsigned long long first_err = 0;
    if ((rc = read_one(c->g_firsterr.as<unsigned long long>(), first_err, s)) != JAM_OK) return rc;
    if ((long long)first_err < n_atoms) {
        long long li = 0;
        unsigned int kind = 0;
        HIP_TRY(hipMemcpyAsync(&li, c->g_atomline.as<long long>() + first_err, sizeof(li), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(&kind, c->g_err.as<unsigned int>() + first_err, sizeof(kind), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        int t = 0;
        while (t + 1 < n && h_rng[t + 1] <= li) t++;
        return fail(JAM_EINVAL, "text " + std::to_string(t) + ", line " + std::to_string(li - h_rng[t] + 1) +
                                    ": failed to parse atom (" +
                                    ((kind & ING_ENONASCII) ? "a non-ASCII character in the first 80 columns"
                                                            : "a numeric field does not parse") +
                                    ")");
    }
    // ---- residue-name index
    HIP_TRY(c->g_keys.reserve(sizeof(uint32_t) * ING_MAX_NAMES * (size_t)std::max(n, 1)));
    HIP_TRY(c->g_nnames.reserve(sizeof(int) * (size_t)std::max(n, 1)));
    if (n > 0)
        hipLaunchKernelGGL(jam::k_ing_names, dim3((unsigned)n), dim3(1024), 0, s, M->atom_off.as<const int>(),
                           M->resname.as<const uint32_t>(), c->g_keys.as<uint32_t>(), c->g_nnames.as<int>());
    HIP_TRY(hipGetLastError());
    std::vector<int> nnames(std::max(n, 1), 0);
    HIP_TRY(hipMemcpyAsync(nnames.data(), c->g_nnames.ptr, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<long long> resoff64(n + 1, 0);
    std::vector<int> resoff(n + 1, 0);
    for (int32_t k = 0; k < n; k++) {
        if (nnames[k] < 0)
            return fail(JAM_ETOOBIG, "text " + std::to_string(k) + " has more than " + std::to_string(ING_MAX_NAMES) +
                                         " distinct residue names");
        resoff64[k + 1] = resoff64[k] + nnames[k];
        resoff[k + 1] = (int)resoff64[k + 1];
    }
    const long long nb = resoff64[n];
    M->n_buckets = nb;
    const size_t nbc = (size_t)std::max<long long>(nb, 1);
    HIP_TRY(upload(M->res_off, resoff, s));
    HIP_TRY(upload(c->g_resoff64, resoff64, s));
    HIP_TRY(M->res_key.reserve(sizeof(uint32_t) * nbc));
    HIP_TRY(M->res_begin.reserve(sizeof(int) * nbc));
    HIP_TRY(M->res_end.reserve(sizeof(int) * nbc));
    HIP_TRY(M->perm.reserve(sizeof(int) * na));
    if (n > 0)
        hipLaunchKernelGGL(jam::k_ing_keys, dim3((unsigned)n), dim3(256), 0, s, c->g_keys.as<const uint32_t>(),
                           c->g_nnames.as<const int>(), c->g_resoff64.as<const long long>(), n, M->res_key.as<uint32_t>());
    HIP_TRY(c->g_bucket.reserve(sizeof(int) * na));
    HIP_TRY(c->g_bcount.reserve(sizeof(int) * nbc));
    HIP_TRY(c->g_bstart.reserve(sizeof(long long) * (nbc + 1)));
    hipLaunchKernelGGL(k_fill_i32, dim3(64), dim3(256), 0, s, c->g_bcount.as<int>(), (long long)nbc, 0);
    if (n_atoms > 0) {
        const unsigned ablocks = (unsigned)((n_atoms + ING_THREADS - 1) / ING_THREADS);
        hipLaunchKernelGGL(jam::k_ing_bucket, dim3(ablocks), dim3(ING_THREADS), 0, s, M->atom_off.as<const int>(), n,
                           n_atoms, M->resname.as<const uint32_t>(), M->res_key.as<const uint32_t>(),
                           c->g_resoff64.as<const long long>(), c->g_bucket.as<int>(), c->g_bcount.as<int>());
        if ((rc = device_scan<int>(c->g_bcount.as<int>(), nb, c->g_bstart.as<long long>(), tmp, 0, s)) != JAM_OK) return rc;
        // atoms sorted by (bucket, serial), stable
        HIP_TRY(c->g_skey.reserve(sizeof(unsigned long long) * 2 * na));
        HIP_TRY(c->g_sval.reserve(sizeof(int) * 2 * na));
        unsigned long long *k0 = c->g_skey.as<unsigned long long>(), *k1 = k0 + na;
        int *v0 = c->g_sval.as<int>(), *v1 = v0 + na;
        hipLaunchKernelGGL(jam::k_ing_sortkeys, dim3(ablocks), dim3(ING_THREADS), 0, s, n_atoms,
                           c->g_bucket.as<const int>(), M->serial.as<const int>(), k0, v0);
        int bucket_bits = 0;
        while ((1ll << bucket_bits) < nb) bucket_bits++;
        int which = 0;
        if ((rc = radix_sort<unsigned long long>(k0, v0, k1, v1, n_atoms, 32 + bucket_bits, c->rs_hist, c->rs_off, tmp,
                                                 s, which)) != JAM_OK)
            return rc;
        hipLaunchKernelGGL(jam::k_ing_perm, dim3(ablocks), dim3(ING_THREADS), 0, s, M->atom_off.as<const int>(), n,
                           n_atoms, (const int *)(which ? v1 : v0), M->perm.as<int>());
        hipLaunchKernelGGL(jam::k_ing_ranges, dim3((unsigned)((nb + ING_THREADS - 1) / ING_THREADS)), dim3(ING_THREADS),
                           0, s, M->atom_off.as<const int>(), c->g_resoff64.as<const long long>(), n, nb,
                           c->g_bstart.as<const long long>(), M->res_begin.as<int>(), M->res_end.as<int>());
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    *out = M.release();
    return JAM_OK;
}

extern "C" int32_t jam_molecules_sizes(const jam_molecules *m, int64_t *atom_off, int64_t *n_buckets)
{
    if (!m) return fail(JAM_EINVAL, "NULL molecules");
    if (n_buckets) *n_buckets = m->n_buckets;
    if (atom_off) {
        std::vector<int> a((size_t)m->n_mols + 1, 0);
        if (m->n_mols > 0) {
            HIP_TRY(hipSetDevice(m->ctx->device));
            HIP_TRY(hipMemcpy(a.data(), m->atom_off.ptr, sizeof(int) * a.size(), hipMemcpyDeviceToHost));
        }
        for (size_t k = 0; k < a.size(); k++) atom_off[k] = a[k];
    }
    return JAM_OK;
}

extern "C" int32_t jam_molecules_copy_atoms(const jam_molecules *m, const jam_atom_fields *f)
{
    if (!m || !f) return fail(JAM_EINVAL, "NULL argument");
    const size_t na = (size_t)m->n_atoms;
    if (na == 0) return JAM_OK;
    const bool extra = f->serial || f->chains || f->altloc || f->icode || f->occupancy || f->tempfactor || f->segid ||
                       f->element || f->charge;
    if (extra && !m->parsed) return fail(JAM_EINVAL, "only molecules parsed on the device keep every atom field");
    HIP_TRY(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    auto get = [&](void *dst, const DBuf &src, size_t bytes) -> hipError_t {
        if (!dst) return hipSuccess;
        return hipMemcpyAsync(dst, src.ptr, bytes, hipMemcpyDeviceToHost, s);
    };
    HIP_TRY(get(f->xyz, m->xyz, sizeof(double) * 3 * na));
    HIP_TRY(get(f->names, m->name, 4 * na));
    HIP_TRY(get(f->resnames, m->resname, 4 * na));
    HIP_TRY(get(f->resseq, m->resseq, sizeof(int) * na));
    HIP_TRY(get(f->chains, m->chains, 2 * na));
    HIP_TRY(get(f->serial, m->serial, sizeof(int) * na));
    HIP_TRY(get(f->altloc, m->altloc, na));
    HIP_TRY(get(f->icode, m->icode, na));
    HIP_TRY(get(f->occupancy, m->occupancy, sizeof(float) * na));
    HIP_TRY(get(f->tempfactor, m->tempfactor, sizeof(float) * na));
    HIP_TRY(get(f->segid, m->segid, 4 * na));
    HIP_TRY(get(f->element, m->element, 2 * na));
    HIP_TRY(get(f->charge, m->charge, sizeof(int) * na));
    HIP_TRY(hipStreamSynchronize(s));
    return JAM_OK;
}

extern "C" int32_t jam_molecules_residue_index(const jam_molecules *m, int32_t *perm, int32_t *res_off,
                                               uint32_t *res_key, int32_t *res_begin, int32_t *res_end)
{
    if (!m) return fail(JAM_EINVAL, "NULL molecules");
    HIP_TRY(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    const size_t na = (size_t)m->n_atoms, nb = (size_t)m->n_buckets;
    if (perm && na) HIP_TRY(hipMemcpyAsync(perm, m->perm.ptr, sizeof(int) * na, hipMemcpyDeviceToHost, s));
    if (res_off) HIP_TRY(hipMemcpyAsync(res_off, m->res_off.ptr, sizeof(int) * ((size_t)m->n_mols + 1),
                                        hipMemcpyDeviceToHost, s));
    if (res_key && nb) HIP_TRY(hipMemcpyAsync(res_key, m->res_key.ptr, 4 * nb, hipMemcpyDeviceToHost, s));
    if (res_begin && nb) HIP_TRY(hipMemcpyAsync(res_begin, m->res_begin.ptr, 4 * nb, hipMemcpyDeviceToHost, s));
    if (res_end && nb) HIP_TRY(hipMemcpyAsync(res_end, m->res_end.ptr, 4 * nb, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return JAM_OK;
}

// ---------------------------------------------------------------------------
// query
// ---------------------------------------------------------------------------

namespace {

int32_t grow_sets(jam_context *c, int64_t cap, int n_sets)
{
    if (n_sets > c->nsets_cap) {
        HIP_TRY(c->cnt.reserve(sizeof(int) * (size_t)n_sets));
        HIP_TRY(c->off.reserve(sizeof(int64_t) * ((size_t)n_sets + 1)));   // + the total
        HIP_TRY(c->kd_bucket.reserve(sizeof(int) * KD_TIERS * (size_t)n_sets));
        HIP_TRY(c->kd_tied.reserve(sizeof(int) * KD_TIERS * (size_t)n_sets));
        c->nsets_cap = n_sets;
    }
    const int64_t blk = cap / JAM_WAVE + n_sets + 1;
    if (cap > c->set_cap) {
        cap = std::max<int64_t>(cap, 1024);
        HIP_TRY(c->list.reserve(sizeof(int) * cap));
        HIP_TRY(c->kidx.reserve(sizeof(int) * cap));
        HIP_TRY(c->l2k.reserve(sizeof(int) * cap));
        HIP_TRY(c->kent.reserve(sizeof(SetEnt) * cap));
        HIP_TRY(c->katom.reserve(sizeof(int) * cap));
        HIP_TRY(c->kscr.reserve((size_t)KD_SCRATCH_BYTES_PER_ENTRY * cap));
        c->set_cap = cap;
    }
    // block and sub-block boxes: set s's boxes start at off[s] / size + s
    const int64_t blk_need = c->set_cap / JAM_WAVE + n_sets + 1;
    if (blk_need > c->blk_cap || blk > c->blk_cap) {
        c->blk_cap = std::max(blk, blk_need);
        HIP_TRY(c->bbox.reserve(sizeof(double) * 6 * (size_t)c->blk_cap));
    }
    const int64_t sblk_need = c->set_cap / JAM_SUB + n_sets + 1;
    if (sblk_need > c->sblk_cap) {
        c->sblk_cap = sblk_need;
        HIP_TRY(c->sbox.reserve(sizeof(double) * 6 * (size_t)c->sblk_cap));
    }
    // per tier: sets, tied sets; then an error word for jam_kd_order
    HIP_TRY(c->kd_count.reserve(sizeof(unsigned int) * (2 * KD_TIERS + 1)));
    return JAM_OK;
}

// Waits for the side-stream work of the last kd launch (before buffers it
// uses are grown, or when a query stops early).
void drain_kd(jam_context *c)
{
    if (!c->kd_pending) return;
    for (int k = 0; k < KD_TIERS; k++) (void)hipStreamSynchronize(c->kd_side[k]);
    c->kd_pending = false;
}

int32_t grow_pairs(jam_context *c, int n_pairs)
{
    if (n_pairs <= c->npairs_cap) return JAM_OK;
    HIP_TRY(c->p_ord.reserve((size_t)JAM_MAX_LEVELS * n_pairs));
    HIP_TRY(c->p_lvl.reserve(sizeof(PairLevel) * (size_t)JAM_MAX_LEVELS * n_pairs));
    HIP_TRY(c->p_nchunk.reserve(sizeof(int) * (size_t)n_pairs));
    HIP_TRY(c->p_base.reserve(sizeof(long long) * ((size_t)n_pairs + 1)));   // + the total
    // [0]: event count, [1, n_pairs]: complete matches per pair, then the kd
    // passes' error word
    HIP_TRY(c->p_cm.reserve(sizeof(long long) * ((size_t)n_pairs + 2)));
    HIP_TRY(c->p_first.reserve(sizeof(long long) * (size_t)n_pairs));
    HIP_TRY(c->p_n1.reserve(sizeof(int) * (size_t)n_pairs));
    c->npairs_cap = n_pairs;
    return JAM_OK;
}

// pinned host words for counter read-backs
int32_t grow_hsmall(jam_context *c, int64_t words)
{
    if (words <= c->hsmall_cap) return JAM_OK;
    words = std::max<int64_t>(words, 1024);
    if (c->hsmall) HIP_TRY(hipHostFree(c->hsmall));
    c->hsmall = nullptr;
    c->hsmall_cap = 0;
    HIP_TRY(hipHostMalloc((void **)&c->hsmall, sizeof(unsigned long long) * (size_t)words, hipHostMallocDefault));
    c->hsmall_cap = words;
    return JAM_OK;
}

int32_t grow_items(jam_context *c, int64_t n_items)
{
    if (n_items <= c->items_cap) return JAM_OK;
    n_items = std::max<int64_t>(n_items, 4096);
    HIP_TRY(c->i_pair.reserve(sizeof(int) * (size_t)n_items));
    c->items_cap = n_items;
    return JAM_OK;
}

// join frontiers (both buffers), rows of `stride` entries
int32_t grow_front(jam_context *c, int64_t rows, int stride)
{
    if (rows <= c->front_cap && stride <= c->front_stride) return JAM_OK;
    rows = std::max<int64_t>(std::max<int64_t>(rows, c->front_cap), 1 << 16);
    stride = std::max(stride, c->front_stride);
    for (int b = 0; b < 2; b++) {
        HIP_TRY(c->j_pair[b].reserve(sizeof(int) * (size_t)rows));
        HIP_TRY(c->j_ent[b].reserve(sizeof(int) * (size_t)rows * stride));
    }
    c->front_cap = rows;
    c->front_stride = stride;
    return JAM_OK;
}

int32_t grow_done(jam_context *c, int64_t rows, int stride)
{
    if (rows <= c->done_cap && stride <= c->done_stride) return JAM_OK;
    rows = std::max<int64_t>(std::max<int64_t>(rows, c->done_cap), 1 << 14);
    stride = std::max(stride, c->done_stride);
    HIP_TRY(c->d_pair.reserve(sizeof(int) * (size_t)rows));
    HIP_TRY(c->d_ent.reserve(sizeof(int) * (size_t)rows * stride));
    c->done_cap = rows;
    c->done_stride = stride;
    return JAM_OK;
}

// keys and permutations of the sort of n matches
int32_t grow_sort(jam_context *c, int64_t n)
{
    if (n <= c->sort_cap) return JAM_OK;
    n = std::max<int64_t>(n, 1 << 14);
    for (int b = 0; b < 2; b++) {
        HIP_TRY(c->s_key[b].reserve(sizeof(unsigned long long) * (size_t)n));
        HIP_TRY(c->s_perm[b].reserve(sizeof(int) * (size_t)n));
        HIP_TRY(c->s_pkey[b].reserve(sizeof(unsigned int) * (size_t)n));
    }
    c->sort_cap = n;
    return JAM_OK;
}

int32_t grow_events(jam_context *c, int64_t cap, int stride)
{
    if (cap > c->event_cap || stride > c->event_stride) {
        cap = std::max<int64_t>(cap, std::max<int64_t>(c->event_cap, 1024));
        stride = std::max(stride, c->event_stride);
        HIP_TRY(c->e_all.reserve(event_layout(cap, stride).bytes));
        c->event_cap = cap;
        c->event_stride = stride;
    }
    return JAM_OK;
}

// clears the kd passes' per-tier counts and their error word
__global__ void k_kd_setup(unsigned int *counts, int n, unsigned int *err)
{
    if ((int)threadIdx.x < n) counts[threadIdx.x] = 0u;
    if (threadIdx.x == 0) *err = 0u;
}

__global__ void k_items_any(DevPairs pr)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pr.n_pairs || *pr.n_items > pr.items_cap) return;   // does not fit: grown and run again
    const int c = pr.nchunk[p];
    const long long b = pr.item_base[p];
    for (int i = 0; i < c; i++) pr.item_pair[b + i] = p;
}

__global__ void k_zero_u64(unsigned long long *v, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        v[i] = 0ull;
}

// workgroups of `kernel` (with `threads` threads) that stay resident on the device
template <class K>
int resident_grid(K kernel, int threads, int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 1024;
    return std::max(1, per_cu) * std::max(1, cus);
}

#define JOIN_KERNEL(M) (jam::k_join<M, JOIN_TW>)

template <int MAXN>
int join_grid(int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, JOIN_KERNEL(MAXN), JOIN_WAVES * JAM_WAVE, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    return std::max(1, per_cu) * std::max(1, cus);
}

// resident workgroups of the k_join variant for rows of `stride` entries
int join_workgroups(jam_context *c, int stride)
{
    const int v = stride <= 8 ? 0 : stride <= 16 ? 1 : stride <= 24 ? 2 : 3;
    if (c->join_grid[v] == 0) {
        c->join_grid[v] = v == 0   ? join_grid<8>(c->device)
                          : v == 1 ? join_grid<16>(c->device)
                          : v == 2 ? join_grid<24>(c->device)
                                   : join_grid<32>(c->device);
        if (c->join_grid[v] == 0) c->join_grid[v] = 1024;
    }
    return c->join_grid[v];
}

// one join level; the grid is the number of workgroups that stay resident
// e0, e1 (profiling, else null): events the runtime records at the
// dispatch's own start and end
void launch_join(jam_context *c, int stride, hipStream_t s, const DevMols &dm, const DevTpls &dt, const DevSets &st,
                 const DevPairs &pr, const DevJoin &jn, const QueryParams &qp, hipEvent_t e0, hipEvent_t e1)
{
    const int v = stride <= 8 ? 0 : stride <= 16 ? 1 : stride <= 24 ? 2 : 3;
    const dim3 g((unsigned)join_workgroups(c, stride)), b(JOIN_WAVES * JAM_WAVE);
    if (v == 0) hipExtLaunchKernelGGL(JOIN_KERNEL(8), g, b, 0, s, e0, e1, 0, dm, dt, st, pr, jn, qp);
    else if (v == 1) hipExtLaunchKernelGGL(JOIN_KERNEL(16), g, b, 0, s, e0, e1, 0, dm, dt, st, pr, jn, qp);
    else if (v == 2) hipExtLaunchKernelGGL(JOIN_KERNEL(24), g, b, 0, s, e0, e1, 0, dm, dt, st, pr, jn, qp);
    else hipExtLaunchKernelGGL(JOIN_KERNEL(32), g, b, 0, s, e0, e1, 0, dm, dt, st, pr, jn, qp);
}


// Launches the kd passes over the candidate sets of `st`: the sets are
// bucketed by size tier, then every tier gets its fast pass on `s` (every
// set: <= 64 entries in one wave's registers, up to KD_BIG with one position
// per thread, above KD_LDS_MAX in global scratch; a set with a tie at a split
// point is listed instead of written) and, on the tier's side stream from the
// end of that fast pass, the exact QuickSelect pass over the listed sets. Sets
// of KD_BIG+1 ... KD_LDS_MAX entries skip the fast pass (at library sizes
// 25-100% of them have a tie) and all go to the exact pass in LDS. With one
// molecule (few sets, latency-bound) the 513-1024 and 257-512 tiers skip the
// fast pass and run the exact pass on all their sets from the start: for a
// handful of big sets it is the kd phase's longest chain, and the fast pass
// would mostly give up on them (with 0.001 A coordinates a big set nearly
// always has equal values at one of its n-1 split points). kd_joined[k] is
// recorded at the end of tier k's passes; every set is written once all are.
// Grids are upper bounds (the counts stay on the device); every kernel loops
// over its own list.
int32_t launch_kd(jam_context *c, const DevMols &dm, const DevTpls &dt, const DevSets &st, hipStream_t s,
                  bool one_molecule)
{
    const int n_sets = st.n_sets;
    unsigned char *scr = c->kscr.as<unsigned char>();
    int *bucket = c->kd_bucket.as<int>();
    int *tied = c->kd_tied.as<int>();
    unsigned int *nb = c->kd_count.as<unsigned int>();
    unsigned int *nt = nb + KD_TIERS;
    const size_t ns = (size_t)n_sets;
    hipLaunchKernelGGL(k_kd_setup, dim3(1), dim3(64), 0, s, nb, 2 * KD_TIERS, st.kd_err);
    hipLaunchKernelGGL(jam::k_kdbucket,
                       dim3(std::max(1, std::min((n_sets + KD_BUCKET_SPAN - 1) / KD_BUCKET_SPAN, 1024))), dim3(256),
                       0, s, st, bucket, nb);
    HIP_TRY(hipGetLastError());
    if (c->kd_grid[0] == 0) {
        const int d = c->device;
        c->kd_grid[0] = resident_grid(jam::k_kdwave, JAM_WAVE, d);
        c->kd_grid[1] = resident_grid(jam::k_kdblock<128>, 128, d);
        c->kd_grid[2] = resident_grid(jam::k_kdblock<256>, 256, d);
        c->kd_grid[3] = resident_grid(jam::k_kdblock<512>, 512, d);
        c->kd_grid[4] = resident_grid(jam::k_kdblock<1024>, 1024, d);
        c->kd_grid[5] = resident_grid(jam::k_kdbig, KD_THREADS, d);
        c->kd_grid[6] = resident_grid(jam::k_kdexact<64>, 64, d);
        c->kd_grid[7] = resident_grid(jam::k_kdexact<128>, 128, d);
        c->kd_grid[8] = resident_grid(jam::k_kdexact<256>, 256, d);
        c->kd_grid[9] = resident_grid(jam::k_kdexact<512>, 512, d);
        c->kd_grid[10] = resident_grid(jam::k_kdexact<1024>, 1024, d);
        c->kd_grid[11] = resident_grid(jam::k_kdbig_exact, KD_EXACT_THREADS, d);
        c->kd_grid[12] = resident_grid(jam::k_kdexact_lds<KD_LDS_SMALL_THREADS, 8>, KD_LDS_SMALL_THREADS, d);
        c->kd_grid[13] = resident_grid(jam::k_kdexact_lds<KD_LDS_MAX / 8, 8>, KD_LDS_MAX / 8, d);
    }
    auto grid = [&](int k) { return dim3((unsigned)std::min(n_sets, c->kd_grid[k])); };
    auto launch_fast = [&](int k, hipStream_t x) {
        const size_t o = (size_t)k * ns;
        switch (k) {
        case 0:
            hipLaunchKernelGGL(jam::k_kdwave, grid(0), dim3(JAM_WAVE), 0, x, dm, dt, st, bucket, nb, tied, nt);
            break;
        case 1:
            hipLaunchKernelGGL((jam::k_kdblock<128>), grid(1), dim3(128), 0, x, dm, dt, st, bucket + o, nb + 1,
                               tied + o, nt + 1);
            break;
        case 2:
            hipLaunchKernelGGL((jam::k_kdblock<256>), grid(2), dim3(256), 0, x, dm, dt, st, bucket + o, nb + 2,
                               tied + o, nt + 2);
            break;
        case 3:
            hipLaunchKernelGGL((jam::k_kdblock<512>), grid(3), dim3(512), 0, x, dm, dt, st, bucket + o, nb + 3,
                               tied + o, nt + 3);
            break;
        case 4:
            hipLaunchKernelGGL((jam::k_kdblock<1024>), grid(4), dim3(1024), 0, x, dm, dt, st, bucket + o, nb + 4,
                               tied + o, nt + 4);
            break;
        default:
            hipLaunchKernelGGL(jam::k_kdbig, grid(5), dim3(KD_THREADS), 0, x, dm, dt, st, scr, bucket + o, nb + 5,
                               tied + o, nt + 5, KD_LDS_MAX);
            break;
        }
    };
    // the exact pass of tier k over the sets its fast pass listed, or
    // (`all`) over every set of the tier
    auto launch_exact = [&](int k, hipStream_t x, bool all) {
        const size_t o = (size_t)k * ns;
        const int *lst = all ? bucket + o : tied + o;
        const unsigned int *lc = all ? nb + k : nt + k;
        switch (k) {
        case 0:
            hipLaunchKernelGGL((jam::k_kdexact<64>), grid(6), dim3(64), 0, x, dm, dt, st, lst, lc);
            break;
        case 1:
            hipLaunchKernelGGL((jam::k_kdexact<128>), grid(7), dim3(128), 0, x, dm, dt, st, lst, lc);
            break;
        case 2:
            hipLaunchKernelGGL((jam::k_kdexact<256>), grid(8), dim3(256), 0, x, dm, dt, st, lst, lc);
            break;
        case 3:
            hipLaunchKernelGGL((jam::k_kdexact<512>), grid(9), dim3(512), 0, x, dm, dt, st, lst, lc);
            break;
        case 4:
            hipLaunchKernelGGL((jam::k_kdexact<1024>), grid(10), dim3(1024), 0, x, dm, dt, st, lst, lc);
            break;
        default:
            // every set of the tier of 8 * KD_LDS_SMALL_THREADS + 1 ...
            // KD_LDS_MAX entries (most of them have a tie) in LDS (the
            // smaller ones run on kd_lds), then the larger sets the fast
            // pass listed in global scratch
            hipLaunchKernelGGL((jam::k_kdexact_lds<KD_LDS_MAX / 8, 8>), grid(13), dim3(KD_LDS_MAX / 8), 0, x, dm, dt,
                               st, bucket + o, nb + 5, KD_LDS_SMALL_THREADS * 8);
            hipLaunchKernelGGL(jam::k_kdbig_exact, grid(11), dim3(KD_EXACT_THREADS), 0, x, dm, dt, st, scr, lst, lc,
                               KD_LDS_MAX);
            break;
        }
    };
    c->kd_pending = true;
    // largest tier first: its exact pass, the longest, starts earliest (with
    // one molecule the exact-only tiers first: sets above 1,024 entries are
    // rare there)
    static const int by_size[KD_TIERS] = {5, 4, 3, 2, 1, 0};
    static const int one_mol[KD_TIERS] = {4, 3, 5, 2, 1, 0};
    static_assert(KD_TIERS == 6, "tier orders list six tiers");
    for (int i = 0; i < KD_TIERS; i++) {
        const int k = one_molecule ? one_mol[i] : by_size[i];
        const bool all = one_molecule && (k == 3 || k == 4);
        if (!all) launch_fast(k, s);
        HIP_TRY(hipEventRecord(c->kd_fast[k], s));
        HIP_TRY(hipStreamWaitEvent(c->kd_side[k], c->kd_fast[k], 0));
        if (k == 5) {
            HIP_TRY(hipStreamWaitEvent(c->kd_lds, c->kd_fast[k], 0));
            hipLaunchKernelGGL((jam::k_kdexact_lds<KD_LDS_SMALL_THREADS, 8>), grid(12), dim3(KD_LDS_SMALL_THREADS), 0,
                               c->kd_lds, dm, dt, st, bucket + (size_t)k * ns, nb + k, KD_BIG);
            HIP_TRY(hipEventRecord(c->kd_lds_done, c->kd_lds));
        }
        launch_exact(k, c->kd_side[k], all);
        if (k == 5) HIP_TRY(hipStreamWaitEvent(c->kd_side[k], c->kd_lds_done, 0));
        HIP_TRY(hipEventRecord(c->kd_joined[k], c->kd_side[k]));
    }
    HIP_TRY(hipGetLastError());
    return JAM_OK;
}

}  // namespace

// One device array of counters, read back in one copy after the join: rows
// per join level, complete matches, the overflow flag, then the sizes phase A
// finds (candidate-set entries, level-0 items, largest set).
constexpr int CNT_DONE = JAM_MAX_LEVELS;
constexpr int CNT_OVF = JAM_MAX_LEVELS + 1;
constexpr int CNT_TOTAL = JAM_MAX_LEVELS + 2;
constexpr int CNT_ITEMS = JAM_MAX_LEVELS + 3;
constexpr int CNT_MAXCNT = JAM_MAX_LEVELS + 4;
constexpr int CNT_WORDS = JAM_MAX_LEVELS + 5;

extern "C" int32_t jam_query(jam_context *c, const jam_templates *tp, const jam_molecules *ms,
                             const jam_params *p, jam_result **out)
{
    if (!c || !tp || !ms || !p || !out) return fail(JAM_EINVAL, "NULL argument");
    if (tp->ctx != c || ms->ctx != c) return fail(JAM_EINVAL, "objects belong to another context");
    if (p->ignore_chain < 0 || p->ignore_chain > 2) return fail(JAM_EINVAL, "invalid ignore_chain");
    *out = nullptr;
    HIP_TRY(hipSetDevice(c->device));
    drain_kd(c);
    std::unique_ptr<jam_result> R(new jam_result());
    R->ctx = c;
    R->gen = ++c->query_gen;
    const int T = tp->n_templates, M = ms->n_mols, U = tp->n_types;
    const int64_t n_pairs64 = (int64_t)T * M;
    if (n_pairs64 > INT32_MAX / 64) return fail(JAM_ETOOBIG, "too many (template, molecule) pairs");
    if ((int64_t)M * U > INT32_MAX / 64) return fail(JAM_ETOOBIG, "too many (molecule, atom type) sets");
    const int n_pairs = (int)n_pairs64;
    R->n_pairs = n_pairs;
    R->stride = tp->stride;
    R->pair_cm.assign(n_pairs, 0);
    if (n_pairs == 0 || U == 0) {
        *out = R.release();
        return JAM_OK;
    }
    const int n_sets = M * U;
    hipStream_t s = c->stream;
    QueryParams qp;
    qp.rmsd_threshold = p->rmsd_threshold;
    qp.distance_cutoff = p->distance_cutoff;
    qp.max_dynamic_distance = p->max_dynamic_distance;
    qp.ignore_chain = p->ignore_chain;
    qp.reorder = p->reorder ? 1 : 0;
    qp.chunk = 1;
    const int stride = tp->stride;

    int32_t rc;
    if ((rc = grow_sets(c, std::max<int64_t>(c->set_cap, 1 << 16), n_sets)) != JAM_OK) return rc;
    if ((rc = grow_events(c, std::max<int64_t>(c->event_cap, 1 << 12), stride)) != JAM_OK) return rc;
    if ((rc = grow_pairs(c, n_pairs)) != JAM_OK) return rc;
    if ((rc = grow_items(c, std::max<int64_t>(c->items_cap, 4096))) != JAM_OK) return rc;
    if ((rc = grow_hsmall(c, std::max<int64_t>(CNT_WORDS, (int64_t)n_pairs + 2))) != JAM_OK) return rc;
    HIP_TRY(c->j_count.reserve(sizeof(unsigned long long) * CNT_WORDS));
    DevMols dm = ms->dev();
    DevTpls dt = tp->dev();
    const int pair_blocks = (n_pairs + 255) / 256;

    for (int attempt = 0; attempt < 8; attempt++) {
        DevSets st;
        st.n_sets = n_sets;
        st.cap = c->set_cap;
        st.blk_cap = c->blk_cap;
        st.cnt = c->cnt.as<int>();
        st.off = c->off.as<int64_t>();
        st.list = c->list.as<int>();
        st.kidx = c->kidx.as<int>();
        st.l2k = c->l2k.as<int>();
        st.kent = c->kent.as<SetEnt>();
        st.katom = c->katom.as<int>();
        st.bbox = c->bbox.as<double>();
        st.sbox = c->sbox.as<double>();
        unsigned long long *cnts = c->j_count.as<unsigned long long>();
        st.total = (int64_t *)(cnts + CNT_TOTAL);
        st.maxcnt = (int *)(cnts + CNT_MAXCNT);
        // the kd error word sits after the per-pair counts, which are read
        // back at the end of the query
        unsigned long long *pcm = (unsigned long long *)c->p_cm.ptr;
        st.kd_err = (unsigned int *)(pcm + (size_t)n_pairs + 1);
        DevPairs pr;
        pr.n_pairs = n_pairs;
        pr.n_items = (const long long *)(cnts + CNT_ITEMS);
        pr.items_cap = c->items_cap;
        pr.ord = c->p_ord.as<unsigned char>();
        pr.lvl = c->p_lvl.as<PairLevel>();
        pr.nchunk = c->p_nchunk.as<int>();
        pr.n1 = c->p_n1.as<int>();
        pr.item_base = c->p_base.as<long long>();
        pr.item_pair = c->i_pair.as<int>();

        const int wpb = 4;  // waves per block of the per-set candidate kernels
        const int set_blocks = (n_sets + wpb - 1) / wpb;
        HIP_TRY(hipEventRecord(c->ev[0], s));
        hipLaunchKernelGGL(jam::k_candidates<false>, dim3(set_blocks), dim3(64 * wpb), 0, s, dm, dt, st, pr,
                           (long long *)nullptr);
        HIP_TRY(hipGetLastError());
        if ((rc = device_scan<int, int64_t>(st.cnt, n_sets, st.off, c->q_scan, 0, s)) != JAM_OK) return rc;
        hipLaunchKernelGGL(k_max_i32, dim3(1), dim3(1024), 0, s, (const int *)st.cnt, (long long)n_sets, st.maxcnt);
        hipLaunchKernelGGL(jam::k_pairs, dim3((unsigned)((n_pairs + 7) / 8)), dim3(256), 0, s, dt, st, pr, qp);
        HIP_TRY(hipGetLastError());
        if ((rc = device_scan<int, long long>(pr.nchunk, n_pairs, pr.item_base, c->q_scan, 0, s)) != JAM_OK)
            return rc;
        // no read-back here: the buffers keep the sizes of earlier queries,
        // every kernel below checks the sizes on the device, and the host
        // reads them with the join's counters, growing and running the query
        // again when something did not fit
        hipLaunchKernelGGL(jam::k_candidates<true>, dim3(set_blocks), dim3(64 * wpb), 0, s, dm, dt, st, pr,
                           (long long *)(cnts + CNT_ITEMS));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(c->ev[1], s));
        if ((rc = launch_kd(c, dm, dt, st, s, M == 1)) != JAM_OK) {
            drain_kd(c);
            return rc;
        }
        // the join's counters and level-0 flags do not depend on the kd
        // order: set them up while the exact passes still run
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, cnts, (long long)CNT_TOTAL);
        hipLaunchKernelGGL(k_items_any, dim3(pair_blocks), dim3(256), 0, s, pr);
        for (int k = 0; k < KD_TIERS; k++) HIP_TRY(hipStreamWaitEvent(s, c->kd_joined[k], 0));
        c->kd_pending = false;   // the rest of the query is ordered after every kd pass
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(c->ev[2], s));

        // ---- the join: one launch per level, back to back; the counters
        // stay on the device until the end
        const int levels = std::max(1, stride - 1);
        if ((rc = grow_front(c, std::max<int64_t>(c->front_cap, 1 << 14), stride)) != JAM_OK) return rc;
        if ((rc = grow_done(c, std::max<int64_t>(c->done_cap, 1 << 14), stride)) != JAM_OK) return rc;
        unsigned long long *fcount = cnts;   // rows per level
        unsigned long long *n_done_dev = cnts + CNT_DONE;
        unsigned int *overflow = (unsigned int *)(cnts + CNT_OVF);
        DevJoin jn{};
        jn.stride = c->front_stride;
        jn.in_stride = 1;
        jn.out_stride = 1;
        jn.n_done = n_done_dev;
        jn.done_cap = c->done_cap;
        jn.done_pair = c->d_pair.as<int>();
        jn.done_ent = c->d_ent.as<int>();
        jn.overflow = overflow;
        if (c->done_stride != c->front_stride) {
            if ((rc = grow_done(c, c->done_cap, c->front_stride)) != JAM_OK) return rc;
            jn.done_pair = c->d_pair.as<int>();
            jn.done_ent = c->d_ent.as<int>();
        }
        {
            jn.level = 0;
            jn.n_out = fcount;
            jn.out_cap = c->front_cap;
            jn.out_pair = c->j_pair[0].as<int>();
            jn.out_ent = c->j_ent[0].as<int>();
            const unsigned l0_blocks = (unsigned)std::min<int64_t>(4096, (c->items_cap + 255) / 256);
            hipLaunchKernelGGL(jam::k_level0, dim3(l0_blocks), dim3(256), 0, s, st, pr, jn);
            HIP_TRY(hipGetLastError());
            for (int L = 1; L <= levels; L++) {
                const int bi = (L - 1) & 1, bo = L & 1;
                jn.level = L;
                jn.n_in = fcount + (L - 1);
                jn.in_cap = c->front_cap;
                jn.in_pair = c->j_pair[bi].as<int>();
                jn.in_ent = c->j_ent[bi].as<int>();
                // the frontier of level L - 1 holds L entries per row (a
                // child of level L exists only for templates of more than
                // L + 1 atoms, so L + 1 never exceeds the row buffers' width)
                jn.in_stride = L;
                jn.out_stride = std::min(L + 1, jn.stride);
                jn.n_out = fcount + L;
                jn.out_cap = c->front_cap;
                jn.out_pair = c->j_pair[bo].as<int>();
                jn.out_ent = c->j_ent[bo].as<int>();
                launch_join(c, stride, s, dm, dt, st, pr, jn, qp, c->profile ? c->xev[2 * (L - 1) + 0] : nullptr,
                            c->profile ? c->xev[2 * (L - 1) + 1] : nullptr);
                HIP_TRY(hipGetLastError());
            }
        }
        unsigned long long *jc = c->hsmall;
        HIP_TRY(hipMemcpyAsync(jc, cnts, sizeof(unsigned long long) * CNT_WORDS, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        const int64_t total = (int64_t)jc[CNT_TOTAL];
        const int64_t n0 = (int64_t)jc[CNT_ITEMS];
        int maxcnt = 0;
        std::memcpy(&maxcnt, jc + CNT_MAXCNT, sizeof(int));
        if (total >= INT32_MAX) return fail(JAM_ETOOBIG, "more than 2^31 candidate-set entries in one query");
        const unsigned int ovf = (unsigned int)(jc[CNT_OVF] & 0xFFFFFFFFull);
        const int64_t n_done = (int64_t)jc[CNT_DONE];
        const bool regrow = total > c->set_cap || n0 > c->items_cap || n0 > c->front_cap;
        if (regrow) {
            if (total > c->set_cap && (rc = grow_sets(c, total, n_sets)) != JAM_OK) return rc;
            if (n0 > c->items_cap && (rc = grow_items(c, n0)) != JAM_OK) return rc;
            if (n0 > c->front_cap && (rc = grow_front(c, std::max<int64_t>(2 * c->front_cap, n0), stride)) != JAM_OK)
                return rc;
            continue;
        }
        unsigned long long rows[JAM_MAX_LEVELS];   // the read-back area is reused below
        std::memcpy(rows, jc, sizeof(rows));
        if (ovf || n_done > c->done_cap) {
            // rows that did not fit: grow to what was seen (deeper levels saw
            // truncated inputs, so this may take another round)
            int64_t need = 0;
            for (int L = 0; L < JAM_MAX_LEVELS; L++) need = std::max<int64_t>(need, (int64_t)jc[L]);
            if ((rc = grow_front(c, std::max<int64_t>(2 * c->front_cap, need), stride)) != JAM_OK) return rc;
            if ((rc = grow_done(c, std::max<int64_t>(c->done_cap, 2 * n_done), c->front_stride)) != JAM_OK)
                return rc;
            continue;
        }
        HIP_TRY(hipEventRecord(c->ev[3], s));

        // ---- reference order of the complete matches: the order of the key
        // (pair, level 0 position, level 1 position, ...), by rank counting
        // when the matches are few, else by LSD radix passes over the level
        // positions (`bits` bits each, `per` levels per 64-bit key), the last
        // levels first, then the pair
        if ((rc = grow_sort(c, n_done)) != JAM_OK) return rc;
        if ((rc = grow_events(c, n_done, stride)) != JAM_OK) return rc;
        int cur = 0;
        const unsigned gs = (unsigned)((n_done + 255) / 256);
        int pbits = 1;
        while ((1ll << pbits) < (long long)n_pairs) pbits++;
        int bits = 1;
        while ((1ll << bits) <= (long long)maxcnt) bits++;
        const int words = (pbits + stride * bits + 63) / 64;
        if (n_done > 0 && n_done <= SORT_RANK_MAX && words <= SORT_RANK_WORDS) {
            // rank counting: same permutation as the radix passes below
            HIP_TRY(c->s_key[0].reserve(sizeof(unsigned long long) * (size_t)n_done * words));
            unsigned long long *fk = c->s_key[0].as<unsigned long long>();
            int *rank = c->s_perm[1].as<int>();
            const dim3 tiles((unsigned)((n_done + 255) / 256), (unsigned)((n_done + SORT_RANK_TJ - 1) / SORT_RANK_TJ));
            switch (words) {
#define JAM_RANK_SORT(W)                                                                                        \
    case W:                                                                                                     \
        hipLaunchKernelGGL(jam::k_fullkey<W>, dim3(gs), dim3(256), 0, s, st, dt, pr, jn, (long long)n_done, stride, \
                           bits, pbits, fk, rank);                                                              \
        hipLaunchKernelGGL(jam::k_rank<W>, tiles, dim3(256), 0, s, (const unsigned long long *)fk,                \
                           (long long)n_done, rank);                                                            \
        break;
                JAM_RANK_SORT(1)
                JAM_RANK_SORT(2)
                JAM_RANK_SORT(3)
                JAM_RANK_SORT(4)
#undef JAM_RANK_SORT
            }
            hipLaunchKernelGGL(jam::k_rank_scatter, dim3(gs), dim3(256), 0, s, (const int *)rank, (long long)n_done,
                               c->s_perm[0].as<int>());
            HIP_TRY(hipGetLastError());
        } else if (n_done > 0) {
            const int per = std::max(1, 64 / bits);
            const int nt = stride;
            // the matches by template size (stably): a match's levels at or
            // past its template's size are 0, so a group of levels starting
            // at l0 needs sorting only over the matches of more than l0
            // atoms, a suffix; that suffix's order is the only one to change,
            // so both permutation buffers keep the rest as it was
            hipLaunchKernelGGL(jam::k_iota, dim3(gs), dim3(256), 0, s, c->s_perm[0].as<int>(), (long long)n_done);
            hipLaunchKernelGGL(jam::k_sizekey, dim3(gs), dim3(256), 0, s, dt, jn, (long long)n_done,
                               c->s_pkey[0].as<unsigned int>());
            {
                int which = 0;
                if ((rc = radix_sort<unsigned int>(c->s_pkey[0].as<unsigned int>(), c->s_perm[0].as<int>(),
                                                   c->s_pkey[1].as<unsigned int>(), c->s_perm[1].as<int>(),
                                                   (long long)n_done, 6, c->rs_hist, c->rs_off, c->q_scan, s,
                                                   which)) != JAM_OK)
                    return rc;
                cur = which;
                HIP_TRY(c->s_starts.reserve(sizeof(long long) * (JAM_MAX_LEVELS + 1)));
                hipLaunchKernelGGL(jam::k_size_starts, dim3(1), dim3(64), 0, s, c->s_pkey[which].as<const unsigned int>(),
                                   (long long)n_done, nt, c->s_starts.as<long long>());
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipMemcpyAsync(c->s_perm[cur ^ 1].as<int>(), c->s_perm[cur].as<const int>(),
                                       sizeof(int) * (size_t)n_done, hipMemcpyDeviceToDevice, s));
                HIP_TRY(hipMemcpyAsync(c->hsmall, c->s_starts.as<const long long>(), sizeof(long long) * (size_t)(nt + 1),
                                       hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
            }
            long long size_le[JAM_MAX_LEVELS + 1];
            std::memcpy(size_le, c->hsmall, sizeof(long long) * (size_t)(nt + 1));
            for (int l0 = ((nt - 1) / per) * per; l0 >= 0; l0 -= per) {
                const int cnt_l = std::min(per, nt - l0);
                const long long a = l0 == 0 ? 0 : size_le[l0];   // matches of at most l0 atoms stay put
                const long long m = (long long)n_done - a;
                if (m <= 1) continue;
                // levels [l0, l0 + cnt_l) packed into cnt_l * bits bits
                hipLaunchKernelGGL(jam::k_sortkey, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, st, dt, pr, jn,
                                   c->s_perm[cur].as<const int>() + a, m, l0, cnt_l, bits,
                                   c->s_key[0].as<unsigned long long>() + a);
                int which = 0;
                if ((rc = radix_sort<unsigned long long>(c->s_key[0].as<unsigned long long>() + a,
                                                         c->s_perm[cur].as<int>() + a,
                                                         c->s_key[1].as<unsigned long long>() + a,
                                                         c->s_perm[cur ^ 1].as<int>() + a, m, cnt_l * bits,
                                                         c->rs_hist, c->rs_off, c->q_scan, s, which)) != JAM_OK)
                    return rc;
                cur ^= which;
            }
            hipLaunchKernelGGL(jam::k_pairkey, dim3(gs), dim3(256), 0, s, jn, T, c->s_perm[cur].as<const int>(),
                               (long long)n_done, c->s_pkey[0].as<unsigned int>());
            int which = 0;
            if ((rc = radix_sort<unsigned int>(c->s_pkey[0].as<unsigned int>(), c->s_perm[cur].as<int>(),
                                               c->s_pkey[1].as<unsigned int>(), c->s_perm[cur ^ 1].as<int>(),
                                               (long long)n_done, pbits, c->rs_hist, c->rs_off, c->q_scan, s,
                                               which)) != JAM_OK)
                return rc;
            cur ^= which;
            HIP_TRY(hipGetLastError());
        }
        const int *perm = c->s_perm[cur].as<const int>();

        // ---- superposition of every complete match; events in that order
        DevEvents ev;
        ev.cap = c->event_cap;
        ev.stride = stride;
        ev.base = c->e_all.as<unsigned char>();
        // the event count, then the complete matches of every pair and the kd
        // error word: read back in one copy
        ev.count = pcm;
        // marks and their scan reuse the (now free) 64-bit sort keys
        int *sel = c->s_key[0].as<int>();
        int *pos = c->s_key[1].as<int>();
        hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min(1024, (n_pairs + 256) / 256)), dim3(256), 0, s, pcm,
                           (long long)n_pairs + 1);
        if (n_done > 0) {
            const long long *first = c->p_first.as<long long>();
            // few matches: the mark pass keeps its superpositions for the emit
            // pass (each is a chain of dependent FP64 work, long next to
            // reading it back)
            double *keep_d = nullptr;
            int *keep_i = nullptr;
            if (n_done <= SORT_RANK_MAX) {
                HIP_TRY(c->keep_d.reserve(sizeof(double) * 16 * (size_t)n_done));
                HIP_TRY(c->keep_i.reserve(sizeof(int) * (size_t)(stride + 2) * (size_t)n_done));
                keep_d = c->keep_d.as<double>();
                keep_i = c->keep_i.as<int>();
            }
            hipLaunchKernelGGL(jam::k_first, dim3(gs), dim3(256), 0, s, jn, perm, (long long)n_done,
                               c->p_first.as<long long>());
            hipLaunchKernelGGL(jam::k_super<false>, dim3(gs), dim3(256), 0, s, dm, dt, st, pr, jn, ev, qp, perm,
                               (long long)n_done, first, sel, (const int *)pos, keep_d, keep_i);
            if ((rc = device_scan<int, int>((const int *)sel, (long long)n_done, pos, c->q_scan, 0, s)) != JAM_OK)
                return rc;
            hipLaunchKernelGGL(jam::k_event_total, dim3(1), dim3(64), 0, s, (const int *)sel, (const int *)pos,
                               (long long)n_done, ev.count);
            hipLaunchKernelGGL(jam::k_super<true>, dim3(gs), dim3(256), 0, s, dm, dt, st, pr, jn, ev, qp, perm,
                               (long long)n_done, first, sel, (const int *)pos, keep_d, keep_i);
            HIP_TRY(hipGetLastError());
        }
        HIP_TRY(hipEventRecord(c->ev[4], s));
        if (n_done > 0) {
            hipLaunchKernelGGL(jam::k_pair_matches, dim3(gs), dim3(256), 0, s, jn, perm, (long long)n_done,
                               (const long long *)c->p_first.as<long long>(), (long long *)(pcm + 1));
            HIP_TRY(hipGetLastError());
        }
        HIP_TRY(hipMemcpyAsync(c->hsmall, pcm, sizeof(unsigned long long) * ((size_t)n_pairs + 2),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        // the error word is the low 32 bits of that slot (only they are
        // cleared and set on the device)
        if ((c->hsmall[(size_t)n_pairs + 1] & 0xFFFFFFFFull) != 0ull)
            return fail(JAM_EDEVICE, "a kd pass reached its round bound (broken QuickSelect invariant)");
        const unsigned long long ne_dev = c->hsmall[0];
        std::memcpy(R->pair_cm.data(), c->hsmall + 1, sizeof(long long) * (size_t)n_pairs);
        if ((int64_t)ne_dev > c->event_cap || (int64_t)ne_dev > n_done)
            return fail(JAM_EDEVICE, "event count exceeds the complete matches");
        float ms_k[5] = {0, 0, 0, 0, 0};
        HIP_TRY(hipEventElapsedTime(&ms_k[0], c->ev[0], c->ev[1]));
        HIP_TRY(hipEventElapsedTime(&ms_k[1], c->ev[1], c->ev[2]));
        HIP_TRY(hipEventElapsedTime(&ms_k[2], c->ev[2], c->ev[3]));
        HIP_TRY(hipEventElapsedTime(&ms_k[3], c->ev[3], c->ev[4]));
        HIP_TRY(hipEventElapsedTime(&ms_k[4], c->ev[0], c->ev[4]));
        for (int k = 0; k < 5; k++) R->ms[k] = ms_k[k];
        const int n_launched = levels;
        R->expand_ms = c->profile ? 0.0 : -1.0;
        for (int q = 0; c->profile && q < n_launched; q++) {
            float x = 0.f;
            HIP_TRY(hipEventElapsedTime(&x, c->xev[2 * q], c->xev[2 * q + 1]));
            R->expand_ms += x;
        }
        R->expand_launches = c->profile ? n_launched : 0;
        for (int q = 0; q < JAM_MAX_LEVELS; q++) R->level_partials[q] = 0;
        for (int L = 1; L < JAM_MAX_LEVELS && L <= n_launched; L++) R->level_partials[L] = (int64_t)rows[L - 1];
        R->n_events = (int64_t)ne_dev;
        HIP_TRY(hipStreamSynchronize(s));
        *out = R.release();
        return JAM_OK;
    }
    return fail(JAM_EDEVICE, "query did not converge after growing its buffers");
}

// ---------------------------------------------------------------------------
// results
// ---------------------------------------------------------------------------

extern "C" int64_t jam_result_n_events(const jam_result *r) { return r ? r->n_events : 0; }

extern "C" int32_t jam_result_stride(const jam_result *r) { return r ? r->stride : 0; }

extern "C" int32_t jam_result_n_pairs(const jam_result *r) { return r ? r->n_pairs : 0; }

namespace {

constexpr size_t JAM_STAGE_BYTES = size_t(64) << 20;

// memcpy spread over a few host threads when it is large
void host_copy(void *dst, const void *src, size_t bytes)
{
    const size_t per = size_t(16) << 20;
    const unsigned nt = (unsigned)std::min<size_t>(8, bytes / per);
    if (nt < 2) {
        std::memcpy(dst, src, bytes);
        return;
    }
    std::vector<std::thread> th;
    const size_t part = (bytes + nt - 1) / nt;
    for (unsigned k = 0; k < nt; k++) {
        const size_t b0 = k * part, b1 = std::min(bytes, b0 + part);
        if (b0 < b1)
            th.emplace_back([=] { std::memcpy((char *)dst + b0, (const char *)src + b0, b1 - b0); });
    }
    for (auto &x : th) x.join();
}

struct CopyPiece {
    void *dst;
    const char *src;
    size_t bytes;
};

}  // namespace

// Copies the fields of n events packed at `base` (device memory, the layout
// of event_layout(n, stride)) to host buffers (any may be NULL) through the
// context's pinned staging halves, on the context's stream.
static int32_t copy_event_fields(jam_context *c, const char *base, size_t n, int stride, int32_t *pair,
                                 int64_t *rank, int32_t *flag, double *rmsd, double *rot, double *c0, double *c1,
                                 int32_t *atoms)
{
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (!c->hstage) {
        HIP_TRY(hipHostMalloc(&c->hstage, 2 * JAM_STAGE_BYTES, hipHostMallocDefault));
        for (auto &ev : c->stage_ev) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    // the fields, cut into pieces of at most one staging half
    // (in the order of the packed layout, so that neighbouring pieces are
    // contiguous on the device)
    std::vector<CopyPiece> pieces;
    const EventLayout L = event_layout((int64_t)n, stride);
    auto add = [&](void *dst, size_t at, size_t bytes) {
        if (!dst) return;
        for (size_t o = 0; o < bytes; o += JAM_STAGE_BYTES)
            pieces.push_back({(char *)dst + o, base + at + o, std::min(JAM_STAGE_BYTES, bytes - o)});
    };
    add(rank, L.rank, sizeof(int64_t) * n);
    add(rmsd, L.rmsd, sizeof(double) * n);
    add(rot, L.rot, sizeof(double) * 9 * n);
    add(c0, L.c0, sizeof(double) * 3 * n);
    add(c1, L.c1, sizeof(double) * 3 * n);
    add(pair, L.pair, sizeof(int32_t) * n);
    add(flag, L.flag, sizeof(int32_t) * n);
    add(atoms, L.atoms, sizeof(int32_t) * (size_t)stride * n);
    // batches of consecutive pieces filling one half each; the copy of batch
    // k + 1 runs while batch k is unpacked from the other half
    struct Batch {
        size_t first, last, half;
    };
    std::vector<Batch> batches;
    for (size_t i = 0; i < pieces.size();) {
        size_t used = 0, j = i;
        while (j < pieces.size() && used + pieces[j].bytes <= JAM_STAGE_BYTES) used += pieces[j++].bytes;
        batches.push_back({i, j, batches.size() & 1});
        i = j;
    }
    auto issue = [&](const Batch &b) -> hipError_t {
        char *h = (char *)c->hstage + b.half * JAM_STAGE_BYTES;
        size_t o = 0;
        for (size_t i = b.first; i < b.last;) {
            // one transfer for every run of pieces contiguous on the device
            size_t j = i + 1, bytes = pieces[i].bytes;
            while (j < b.last && pieces[j].src == pieces[j - 1].src + pieces[j - 1].bytes) bytes += pieces[j++].bytes;
            hipError_t e = hipMemcpyAsync(h + o, pieces[i].src, bytes, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) return e;
            o += bytes;
            i = j;
        }
        return hipEventRecord(c->stage_ev[b.half], s);
    };
    if (!batches.empty()) HIP_TRY(issue(batches[0]));
    for (size_t k = 0; k < batches.size(); k++) {
        if (k + 1 < batches.size()) HIP_TRY(issue(batches[k + 1]));
        const Batch &b = batches[k];
        HIP_TRY(hipEventSynchronize(c->stage_ev[b.half]));
        const char *h = (const char *)c->hstage + b.half * JAM_STAGE_BYTES;
        size_t o = 0;
        for (size_t i = b.first; i < b.last; i++) {
            host_copy(pieces[i].dst, h + o, pieces[i].bytes);
            o += pieces[i].bytes;
        }
    }
    return JAM_OK;
}

extern "C" int32_t jam_result_copy(const jam_result *r, int32_t *pair, int64_t *rank, int32_t *flag,
                                   double *rmsd, double *rot, double *c0, double *c1, int32_t *atoms)
{
    if (!r) return fail(JAM_EINVAL, "NULL result");
    const size_t n = (size_t)r->n_events;
    if (n == 0) return JAM_OK;
    jam_context *c = r->ctx;
    if (!c || c->query_gen != r->gen)
        return fail(JAM_ESTALE, "the events of this result were overwritten by a later query on its context");
    return copy_event_fields(c, c->e_all.as<const char>(), n, r->stride, pair, rank, flag, rmsd, rot, c0, c1,
                             atoms);
}

// events kept in device memory of their own (jam_result_keep)
struct jam_events {
    jam_context *ctx = nullptr;
    int64_t n = 0;
    int stride = 1;
    void *dev = nullptr;
};

extern "C" int32_t jam_result_keep(const jam_result *r, jam_events **out)
{
    if (!r || !out) return fail(JAM_EINVAL, "NULL argument");
    *out = nullptr;
    jam_context *c = r->ctx;
    if (!c || c->query_gen != r->gen)
        return fail(JAM_ESTALE, "the events of this result were overwritten by a later query on its context");
    HIP_TRY(hipSetDevice(c->device));
    std::unique_ptr<jam_events> e(new jam_events());
    e->ctx = c;
    e->n = r->n_events;
    e->stride = r->stride;
    const size_t bytes = event_layout(e->n, e->stride).bytes;
    if (bytes > 0) {
        HIP_TRY(hipMalloc(&e->dev, bytes));
        // ordered before the context's next query, which overwrites the source
        const hipError_t err = hipMemcpyAsync(e->dev, c->e_all.ptr, bytes, hipMemcpyDeviceToDevice, c->stream);
        if (err != hipSuccess) {
            (void)hipFree(e->dev);
            return fail(JAM_EDEVICE, std::string("hipMemcpyAsync: ") + hipGetErrorString(err));
        }
    }
    *out = e.release();
    return JAM_OK;
}

extern "C" int32_t jam_events_copy(const jam_events *e, int32_t *pair, int64_t *rank, int32_t *flag, double *rmsd,
                                   double *rot, double *c0, double *c1, int32_t *atoms)
{
    if (!e) return fail(JAM_EINVAL, "NULL events");
    if (e->n == 0) return JAM_OK;
    return copy_event_fields(e->ctx, (const char *)e->dev, (size_t)e->n, e->stride, pair, rank, flag, rmsd, rot, c0,
                             c1, atoms);
}

extern "C" void jam_events_free(jam_events *e)
{
    if (!e) return;
    (void)hipSetDevice(e->ctx->device);
    if (e->dev) (void)hipFree(e->dev);
    delete e;
}

extern "C" int32_t jam_result_pair_matches(const jam_result *r, int64_t *counts)
{
    if (!r || !counts) return fail(JAM_EINVAL, "NULL argument");
    std::memcpy(counts, r->pair_cm.data(), sizeof(int64_t) * r->pair_cm.size());
    return JAM_OK;
}

extern "C" int32_t jam_result_pair_digests(const jam_result *r, int32_t mol_base, int32_t n_templates,
                                           int64_t *hit_counts, uint64_t *digests)
{
    if (!r || !hit_counts || !digests) return fail(JAM_EINVAL, "NULL argument");
    if (n_templates < 1 || r->n_pairs % n_templates != 0) return fail(JAM_EINVAL, "invalid template count");
    jam_context *c = r->ctx;
    if (!c || c->query_gen != r->gen)
        return fail(JAM_ESTALE, "the events of this result were overwritten by a later query on its context");
    const size_t np = (size_t)r->n_pairs;
    if (np == 0) return JAM_OK;
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    HIP_TRY(c->dig.reserve(2 * sizeof(unsigned long long) * np));
    unsigned long long *d = c->dig.as<unsigned long long>();
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<size_t>(1024, (2 * np + 255) / 256)), dim3(256), 0, s, d,
                       (long long)(2 * np));
    if (r->n_events > 0)
        hipLaunchKernelGGL(jam::k_pair_digest, dim3((unsigned)((r->n_events + 255) / 256)), dim3(256), 0, s,
                           c->e_all.as<const unsigned char>(), (long long)r->n_events, r->stride, n_templates,
                           mol_base, d, d + np);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(hit_counts, d, sizeof(int64_t) * np, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(digests, d + np, sizeof(uint64_t) * np, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return JAM_OK;
}

extern "C" int32_t jam_result_timings(const jam_result *r, double *ms5)
{
    if (!r || !ms5) return fail(JAM_EINVAL, "NULL argument");
    std::memcpy(ms5, r->ms, sizeof(r->ms));
    return JAM_OK;
}

extern "C" int32_t jam_result_level_partials(const jam_result *r, int64_t *counts)
{
    if (!r || !counts) return fail(JAM_EINVAL, "NULL argument");
    std::memcpy(counts, r->level_partials, sizeof(r->level_partials));
    return JAM_OK;
}

extern "C" int32_t jam_context_set_profiling(jam_context *c, int32_t on)
{
    if (!c) return fail(JAM_EINVAL, "NULL context");
    c->profile = on != 0;
    return JAM_OK;
}

extern "C" int32_t jam_result_join_kernels(const jam_result *r, double *ms, int32_t *launches)
{
    if (!r || !ms || !launches) return fail(JAM_EINVAL, "NULL argument");
    *ms = r->expand_ms;
    *launches = r->expand_launches;
    return JAM_OK;
}

extern "C" void jam_result_free(jam_result *r) { delete r; }

// ---------------------------------------------------------------------------
// standalone superposition
// ---------------------------------------------------------------------------

namespace {

// k_superpose_batch over device arrays, on stream s
int32_t superpose_launch(hipStream_t s, const double *da, const double *db, int k, int64_t n_sets, double *drmsd,
                         double *drot, double *dc0, double *dc1)
{
    const int tb = 256;
    hipLaunchKernelGGL(jam::k_superpose_batch, dim3((unsigned)((n_sets + tb - 1) / tb)), dim3(tb), 0, s, da, db, k,
                       (long long)n_sets, drmsd, drot, dc0, dc1);
    HIP_TRY(hipGetLastError());
    return JAM_OK;
}

// copies the outputs back and waits for them
int32_t superpose_fetch(hipStream_t s, int64_t n_sets, const void *drmsd, const void *drot, const void *dc0,
                        const void *dc1, double *rmsd, double *rot, double *c0, double *c1)
{
    HIP_TRY(hipMemcpyAsync(rmsd, drmsd, sizeof(double) * n_sets, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(rot, drot, sizeof(double) * 9 * n_sets, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(c0, dc0, sizeof(double) * 3 * n_sets, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(c1, dc1, sizeof(double) * 3 * n_sets, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return JAM_OK;
}

}  // namespace

extern "C" int32_t jam_superpose_batch(jam_context *c, const double *a, const double *b, int32_t k,
                                       int64_t n_sets, double *rmsd, double *rot, double *c0, double *c1)
{
    if (!c || !a || !b || !rmsd || !rot || !c0 || !c1) return fail(JAM_EINVAL, "NULL argument");
    if (k < 1 || n_sets < 0) return fail(JAM_EINVAL, "invalid sizes");
    if (n_sets == 0) return JAM_OK;
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const size_t pts = (size_t)n_sets * k * 3;
    HIP_TRY(c->sa.reserve(sizeof(double) * pts));
    HIP_TRY(c->sb.reserve(sizeof(double) * pts));
    HIP_TRY(c->srmsd.reserve(sizeof(double) * n_sets));
    HIP_TRY(c->srot.reserve(sizeof(double) * 9 * n_sets));
    HIP_TRY(c->sc0.reserve(sizeof(double) * 3 * n_sets));
    HIP_TRY(c->sc1.reserve(sizeof(double) * 3 * n_sets));
    HIP_TRY(hipMemcpyAsync(c->sa.ptr, a, sizeof(double) * pts, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->sb.ptr, b, sizeof(double) * pts, hipMemcpyHostToDevice, s));
    int32_t rc = superpose_launch(s, c->sa.as<double>(), c->sb.as<double>(), k, n_sets, c->srmsd.as<double>(),
                                  c->srot.as<double>(), c->sc0.as<double>(), c->sc1.as<double>());
    if (rc != JAM_OK) return rc;
    return superpose_fetch(s, n_sets, c->srmsd.ptr, c->srot.ptr, c->sc0.ptr, c->sc1.ptr, rmsd, rot, c0, c1);
}

extern "C" int32_t jam_superpose_batch_noalloc(jam_context *c, const double *a, const double *b, int32_t k,
                                               int64_t n_sets, double *rmsd, double *rot, double *c0, double *c1)
{
    if (!c || !a || !b || !rmsd || !rot || !c0 || !c1) return fail(JAM_EINVAL, "NULL argument");
    if (k < 1 || n_sets < 0) return fail(JAM_EINVAL, "invalid sizes");
    if (n_sets == 0) return JAM_OK;
    // the array, cut into 256-byte aligned pieces: a, b, rmsd, rotations, centroids
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t per_set = sizeof(double) * (6 * (size_t)k + 16);
    if ((size_t)n_sets > JAM_NOALLOC_SCRATCH_BYTES / per_set)
        return fail(JAM_ETOOBIG, "batch larger than the scratch of jam_superpose_batch_noalloc");
    const size_t pts = (size_t)n_sets * k * 3;
    const size_t o_a = 0, o_b = o_a + up(sizeof(double) * pts), o_r = o_b + up(sizeof(double) * pts);
    const size_t o_m = o_r + up(sizeof(double) * n_sets), o_c0 = o_m + up(sizeof(double) * 9 * n_sets);
    const size_t o_c1 = o_c0 + up(sizeof(double) * 3 * n_sets), end = o_c1 + up(sizeof(double) * 3 * n_sets);
    if (end > JAM_NOALLOC_SCRATCH_BYTES)
        return fail(JAM_ETOOBIG, "batch larger than the scratch of jam_superpose_batch_noalloc");
    if (c->device < 0 || c->device >= JAM_MAX_DEVICES) return fail(JAM_EDEVICE, "device ordinal has no scratch");
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_init);
        base = g_scratch[c->device];
    }
    if (!base) return fail(JAM_EDEVICE, "the device's scratch address is unknown");
    std::lock_guard<std::mutex> busy(g_scratch_busy[c->device]);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    unsigned char *w = static_cast<unsigned char *>(base);
    double *da = (double *)(w + o_a), *db = (double *)(w + o_b), *dr = (double *)(w + o_r);
    double *dm = (double *)(w + o_m), *d0 = (double *)(w + o_c0), *d1 = (double *)(w + o_c1);
    HIP_TRY(hipMemcpyAsync(da, a, sizeof(double) * pts, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(db, b, sizeof(double) * pts, hipMemcpyHostToDevice, s));
    int32_t rc = superpose_launch(s, da, db, k, n_sets, dr, dm, d0, d1);
    if (rc != JAM_OK) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    return superpose_fetch(s, n_sets, dr, dm, d0, d1, rmsd, rot, c0, c1);
}

// ---------------------------------------------------------------------------
// standalone kd-tree leaf order
// ---------------------------------------------------------------------------

extern "C" int32_t jam_kd_order(jam_context *c, const double *xyz, const int64_t *off, int32_t n_sets,
                                int32_t *order)
{
    if (!c || !off || (n_sets > 0 && (!xyz || !order))) return fail(JAM_EINVAL, "NULL argument");
    if (n_sets < 0) return fail(JAM_EINVAL, "negative set count");
    if (off[0] != 0) return fail(JAM_EINVAL, "off[0] must be 0");
    for (int32_t k = 0; k < n_sets; k++)
        if (off[k + 1] < off[k]) return fail(JAM_EINVAL, "set offsets decrease");
    const int64_t total = off[n_sets];
    if (total >= INT32_MAX) return fail(JAM_ETOOBIG, "more than 2^31 points");
    if (n_sets == 0 || total == 0) return JAM_OK;
    HIP_TRY(hipSetDevice(c->device));
    drain_kd(c);
    hipStream_t s = c->stream;
    // every set is a molecule of its own whose candidate list is its points
    // in input order; chains and residue numbers are zero
    std::vector<int> aoff(n_sets + 1), cnt(n_sets), list((size_t)total);
    for (int32_t k = 0; k <= n_sets; k++) aoff[k] = (int)off[k];
    for (int32_t k = 0; k < n_sets; k++) {
        cnt[k] = (int)(off[k + 1] - off[k]);
        for (int i = 0; i < cnt[k]; i++) list[(size_t)off[k] + i] = i;
    }
    int32_t rc;
    if ((rc = grow_sets(c, std::max<int64_t>(total, c->set_cap), n_sets)) != JAM_OK) return rc;
    HIP_TRY(c->j_count.reserve(sizeof(unsigned long long) * CNT_WORDS));
    HIP_TRY(c->o_xyz.reserve(sizeof(double) * 3 * (size_t)total));
    HIP_TRY(c->o_zero.reserve(sizeof(int) * (size_t)total));
    HIP_TRY(upload(c->o_aoff, aoff, s));
    HIP_TRY(upload(c->cnt, cnt, s));
    HIP_TRY(upload(c->list, list, s));
    HIP_TRY(hipMemcpyAsync(c->off.ptr, off, sizeof(int64_t) * (size_t)n_sets, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->o_xyz.ptr, xyz, sizeof(double) * 3 * (size_t)total, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(c->o_zero.ptr, 0, sizeof(int) * (size_t)total, s));
    unsigned long long *cnts = c->j_count.as<unsigned long long>();
    const unsigned long long tot = (unsigned long long)total;
    HIP_TRY(hipMemcpyAsync(cnts + CNT_TOTAL, &tot, sizeof(tot), hipMemcpyHostToDevice, s));
    DevMols dm{};
    dm.n_mols = n_sets;
    dm.atom_off = c->o_aoff.as<int>();
    dm.xyz = c->o_xyz.as<double>();
    dm.chain = c->o_zero.as<int>();
    dm.resseq = c->o_zero.as<int>();
    DevTpls dt{};
    dt.n_types = 1;
    DevSets st{};
    st.n_sets = n_sets;
    st.cap = c->set_cap;
    st.blk_cap = c->blk_cap;
    st.cnt = c->cnt.as<int>();
    st.off = c->off.as<int64_t>();
    st.list = c->list.as<int>();
    st.kidx = c->kidx.as<int>();
    st.l2k = c->l2k.as<int>();
    st.kent = c->kent.as<SetEnt>();
    st.katom = c->katom.as<int>();
    st.bbox = c->bbox.as<double>();
    st.sbox = c->sbox.as<double>();
    st.total = (int64_t *)(cnts + CNT_TOTAL);
    st.maxcnt = (int *)(cnts + CNT_MAXCNT);
    st.kd_err = c->kd_count.as<unsigned int>() + 2 * KD_TIERS;
    if ((rc = launch_kd(c, dm, dt, st, s, n_sets == 1)) != JAM_OK) {
        drain_kd(c);
        return rc;
    }
    for (int k = 0; k < KD_TIERS; k++) HIP_TRY(hipStreamWaitEvent(s, c->kd_joined[k], 0));
    std::vector<int> kidx((size_t)total);
    unsigned int err = 0;
    HIP_TRY(hipMemcpyAsync(kidx.data(), st.kidx, sizeof(int) * (size_t)total, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&err, st.kd_err, sizeof(err), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    c->kd_pending = false;
    if (err) return fail(JAM_EDEVICE, "a kd pass reached its round bound (broken QuickSelect invariant)");
    // DFS position p (right child first) holds list element kidx[p] and is
    // leaf n-1-p
    for (int32_t k = 0; k < n_sets; k++) {
        const int64_t o = off[k];
        const int n = cnt[k];
        for (int p = 0; p < n; p++) order[(size_t)o + (n - 1 - p)] = kidx[(size_t)o + p];
    }
    return JAM_OK;
}
