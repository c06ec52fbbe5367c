// neighbor.hpp -- FP64 neighbour finalize, neighbour CSR and hard clusters.
#pragma once
#include "engine.hpp"

namespace gplum {

struct Scratch {
    DArray<unsigned char> tmp;     // hipcub temporary storage
};

struct NbParams {
    double eps2, eps2_sun, m_sun, dt_tree, dt_min, R_search2, R_search3, eta_0, eta_sun0, alpha2;
    Cutoff cut;
};

// Directed neighbour graph after the FP64 decision: row i lists, in ascending
// index, the j that i accepted (NeighborList::n_list, src/neighbor.h:247-250).
struct NbState {
    DArray<uint64_t> keys_in, keys;
    DArray<unsigned char> accepted;
    DArray<int32_t> row_off, number_fp32, nbr_off, nbr_idx;
    DArray<int32_t> info;               // device: candidate-total check, accepted total
    DArray<double> acc_d_new;
    HostBuf info_host;                  // info read back without a wait (nb_resolve)
    bool info_pending = false;
    int n = 0, n_nbr = 0;
};
// the last finalize's check and accepted total, once the stream has passed it
int nb_resolve(NbState &nb);

// Hard clusters: particles with neighbours sorted by (id_cluster, index)
// (HardSystem::makeList, src/hard.h:321-438).
struct ClusterList {
    DArray<int32_t> parent, cflag, cscan, idx_sel, index, index_tmp, sflag, sscan, off;
    DArray<uint64_t> ckey, key_sel, key_sorted;
    DArray<int32_t> size_key, size_key_sorted, order_in, order, max_dev;   // max_dev: cluster count, largest
    HostBuf host;                                                            // read-back of max_dev
    int m = 0, ncl = 0, max_size = 0;
};

NbParams make_nb_params(const gplum_params &p);

// Candidate sort keys i 2^kb + j: kb = cand_key_bits(n) bits hold a particle index.
int cand_key_bits(int n);
// keys of d_cand[0..ncand) sorted ascending (radix sort of their 2 kb bits) into keys_out
int cand_sort(const int2 *d_cand, int ncand, int kb, DArray<uint64_t> &keys_in, uint64_t *keys_out, Scratch &s,
              hipStream_t st);

// correctForceLong[Initial] over the candidate pairs d_cand[0..ncand) emitted by the
// FP32 EP-EP kernel; ps.number holds the FP32 count on entry and the FP64 count on
// exit.  Fills nb (accepted-neighbour CSR).  presorted: nb.keys[0..ncand) already holds
// the sorted keys (d_cand is not read).
int neighbor_finalize(const gplum_params &p, ParticleSoA &ps, int n, const int2 *d_cand, int ncand,
                      bool initial, NbState &nb, Scratch &s, hipStream_t st, bool presorted = false);

// createNeighborCluster + makeList: id_cluster = min id of the component; clusters
// of the particles with neighbours (members by ascending id), offsets cl.off[0..ncl],
// processing order largest first.
// Clusters (connected components of the neighbour graph), members by ascending id, and
// their processing order for the Hermite kernels (estimated cost, costliest first).
int make_clusters(ParticleSoA &ps, int n, NbState &nb, ClusterList &cl, Scratch &s, double dt_tree,
                  hipStream_t st);

}  // namespace gplum
