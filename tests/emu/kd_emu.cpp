// Runs the library's kd kernels (pyjess_amd/csrc/jam_kd.h) on the CPU under
// the HIP emulation of hip/hip_runtime.h, in the order the library launches
// them (launch_kd in jam_api.hip: bucketing by size tier, then per tier the
// fast pass and the exact pass over the sets it found tied, or the exact
// pass over every set of the 513-1024 and 257-512 tiers for one molecule;
// sets of 1,025-8,192 entries go to the LDS exact pass without a fast pass),
// and writes each set's kd leaf order. Test infrastructure only.
//
//   kd_emu INPUT OUTPUT
// INPUT: int32 n_sets, int32 one_molecule, int64 off[n_sets + 1],
//        double xyz[3 * off[n_sets]]
// OUTPUT: int32 order[off[n_sets]] -- per set, its points in leaf order
//         (left subtree first), set-local indices -- then int32 kd_err.
// Every buffer the kernels may read before writing is filled with 0xFF bytes
// first, so that such a read gives garbage instead of zeros.

#include <cstdio>
#include <cstring>
#include <vector>

#include "jam_kd.h"

using namespace jam;

template <class T>
static T *poisoned(std::vector<T> &v, size_t n)
{
    v.assign(n ? n : 1, T());
    std::memset((void *)v.data(), 0xFF, sizeof(T) * v.size());
    return v.data();
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: kd_emu INPUT OUTPUT\n");
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t n_sets = 0, one_molecule = 0;
    if (std::fread(&n_sets, 4, 1, f) != 1 || std::fread(&one_molecule, 4, 1, f) != 1) return 2;
    std::vector<int64_t> off(n_sets + 1);
    if (std::fread(off.data(), 8, off.size(), f) != off.size()) return 2;
    const int64_t total = off[n_sets];
    std::vector<double> xyz(3 * (size_t)total + 3);
    if (std::fread(xyz.data(), 8, 3 * (size_t)total, f) != 3 * (size_t)total) return 2;
    std::fclose(f);

    // one molecule per set, one atom type: set s is molecule s, its candidate
    // list is its atoms in input order
    std::vector<int> atom_off(n_sets + 1), chain(total + 1, 0), resseq(total + 1, 0);
    for (int s = 0; s <= n_sets; s++) atom_off[s] = (int)off[s];
    DevMols ms{};
    ms.n_mols = n_sets;
    ms.atom_off = atom_off.data();
    ms.xyz = xyz.data();
    ms.chain = chain.data();
    ms.resseq = resseq.data();
    DevTpls tp{};
    tp.n_types = 1;

    std::vector<int> cnt(n_sets), list(total + 1);
    std::vector<int64_t> soff(n_sets + 1);
    int maxcnt = 0;
    for (int s = 0; s < n_sets; s++) {
        cnt[s] = (int)(off[s + 1] - off[s]);
        soff[s] = off[s];
        maxcnt = std::max(maxcnt, cnt[s]);
        for (int e = 0; e < cnt[s]; e++) list[off[s] + e] = e;
    }
    soff[n_sets] = total;
    std::vector<int> kidx, l2k, katom, bucket, tied;
    std::vector<SetEnt> kent;
    std::vector<double> bbox, sbox;
    std::vector<unsigned char> scr;
    const int64_t cap = std::max<int64_t>(total, 1);
    const int64_t blk_cap = cap / JAM_WAVE + n_sets + 1;
    int64_t total_dev = total;
    unsigned int kd_err = 0;
    DevSets st{};
    st.n_sets = n_sets;
    st.cap = cap;
    st.blk_cap = blk_cap;
    st.cnt = cnt.data();
    st.off = soff.data();
    st.list = list.data();
    st.kidx = poisoned(kidx, cap);
    st.l2k = poisoned(l2k, cap);
    st.kent = poisoned(kent, cap);
    st.katom = poisoned(katom, cap);
    st.bbox = poisoned(bbox, 6 * (size_t)blk_cap);
    st.sbox = poisoned(sbox, 6 * (size_t)(cap / JAM_SUB + n_sets + 1));
    st.total = &total_dev;
    st.maxcnt = &maxcnt;
    st.kd_err = &kd_err;
    unsigned char *scratch = poisoned(scr, (size_t)KD_SCRATCH_BYTES_PER_ENTRY * cap);
    int *bk = poisoned(bucket, (size_t)KD_TIERS * n_sets);
    int *td = poisoned(tied, (size_t)KD_TIERS * n_sets);
    unsigned int counts[2 * KD_TIERS] = {};
    unsigned int *nb = counts, *nt = counts + KD_TIERS;

    // a small grid: the kernels stride over their lists of sets
    const dim3 g(2);
    emu::launch(k_kdbucket, dim3(1), dim3(256), st, bk, nb);
    static const int by_size[KD_TIERS] = {5, 4, 3, 2, 1, 0};
    static const int one_mol[KD_TIERS] = {4, 3, 5, 2, 1, 0};
    for (int i = 0; i < KD_TIERS; i++) {
        const int k = one_molecule ? one_mol[i] : by_size[i];
        const bool all = one_molecule && (k == 3 || k == 4);
        const size_t o = (size_t)k * n_sets;
        if (!all) {
            switch (k) {
            case 0: emu::launch(k_kdwave, g, dim3(JAM_WAVE), ms, tp, st, (const int *)bk, (const unsigned *)nb, td, nt); break;
            case 1: emu::launch(k_kdblock<128>, g, dim3(128), ms, tp, st, (const int *)bk + o, (const unsigned *)nb + 1, td + o, nt + 1); break;
            case 2: emu::launch(k_kdblock<256>, g, dim3(256), ms, tp, st, (const int *)bk + o, (const unsigned *)nb + 2, td + o, nt + 2); break;
            case 3: emu::launch(k_kdblock<512>, g, dim3(512), ms, tp, st, (const int *)bk + o, (const unsigned *)nb + 3, td + o, nt + 3); break;
            case 4: emu::launch(k_kdblock<1024>, g, dim3(1024), ms, tp, st, (const int *)bk + o, (const unsigned *)nb + 4, td + o, nt + 4); break;
            default: emu::launch(k_kdbig, g, dim3(KD_THREADS), ms, tp, st, scratch, (const int *)bk + o, (const unsigned *)nb + 5, td + o, nt + 5, KD_LDS_MAX); break;
            }
        }
        const int *lst = all ? bk + o : td + o;
        const unsigned *lc = all ? nb + k : nt + k;
        switch (k) {
        case 0: emu::launch(k_kdexact<64>, g, dim3(64), ms, tp, st, lst, lc); break;
        case 1: emu::launch(k_kdexact<128>, g, dim3(128), ms, tp, st, lst, lc); break;
        case 2: emu::launch(k_kdexact<256>, g, dim3(256), ms, tp, st, lst, lc); break;
        case 3: emu::launch(k_kdexact<512>, g, dim3(512), ms, tp, st, lst, lc); break;
        case 4: emu::launch(k_kdexact<1024>, g, dim3(1024), ms, tp, st, lst, lc); break;
        default:
            emu::launch(k_kdexact_lds<KD_LDS_SMALL_THREADS, 8>, g, dim3(KD_LDS_SMALL_THREADS), ms, tp, st, (const int *)bk + o, (const unsigned *)nb + 5, KD_BIG);
            emu::launch(k_kdexact_lds<KD_LDS_MAX / 8, 8>, g, dim3(KD_LDS_MAX / 8), ms, tp, st, (const int *)bk + o, (const unsigned *)nb + 5, KD_LDS_SMALL_THREADS * 8);
            emu::launch(k_kdbig_exact, g, dim3(KD_EXACT_THREADS), ms, tp, st, scratch, lst, lc, KD_LDS_MAX);
            break;
        }
    }

    // kd DFS position p (right child first) holds list element kidx[p] and is
    // leaf n - 1 - p
    std::vector<int32_t> order(total + 1, -1);
    for (int s = 0; s < n_sets; s++) {
        const int n = cnt[s];
        for (int j = 0; j < n; j++) order[off[s] + j] = kidx[off[s] + n - 1 - j];
    }
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(order.data(), 4, (size_t)total, o);
    const int32_t err = (int32_t)kd_err;
    std::fwrite(&err, 4, 1, o);
    std::fclose(o);
    return 0;
}
