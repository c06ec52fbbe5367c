// engine.hpp -- device-resident P3T system (particles in HBM, SoA, Morton order).
#pragma once
#include <vector>

#include "soft.hpp"

namespace gplum {

// Device array with grow-only capacity.
template <class T>
struct DArray {
    T *p = nullptr;
    size_t n = 0, cap = 0;
    DArray() = default;
    DArray(const DArray &) = delete;
    DArray &operator=(const DArray &) = delete;
    ~DArray() {
        if (p) (void)hipFree(p);
    }
    hipError_t ensure(size_t count) {
        n = count;
        if (count <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        size_t want = count + count / 4 + 64;
        hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
        if (e == hipSuccess) cap = want;
        else cap = 0;
        return e;
    }
    T *get() const { return p; }
    void swap(DArray &o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(cap, o.cap);
    }
};

#define GPLUM_ENSURE(arr, cnt) GPLUM_HIP((arr).ensure(cnt))

// Pinned host memory for the small device-to-host reads of a step (grown on demand, kept
// until destroyed): a copy into it is one DMA transfer, where a copy into pageable memory
// goes through the runtime's staging.  Every user synchronises before reading it.
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 4 + 256;
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) cap = want;
        return e;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// Particle state, structure of arrays.  Vectors are stored interleaved (x,y,z).
struct ParticleSoA {
    DArray<int64_t> id, id_cluster;
    DArray<double> mass, r_planet, f;
    DArray<double> pos, vel, acc, acc_d, jerk_d, acc_s, jerk_s;   // 3 per particle
    DArray<double> phi, phi_d, phi_s, r_out, r_out_inv, r_search, v_disp, time, dt, acc0;
    DArray<int32_t> number, rank, id_max, id_min, n_cluster, flags;
    DArray<float> fc_rate;      // forecast neighbour-candidate rate (travels with the particle)
    int ensure(int n);
    void swap(ParticleSoA &o) {
        id.swap(o.id); id_cluster.swap(o.id_cluster);
        mass.swap(o.mass); r_planet.swap(o.r_planet); f.swap(o.f);
        pos.swap(o.pos); vel.swap(o.vel); acc.swap(o.acc); acc_d.swap(o.acc_d);
        jerk_d.swap(o.jerk_d); acc_s.swap(o.acc_s); jerk_s.swap(o.jerk_s);
        phi.swap(o.phi); phi_d.swap(o.phi_d); phi_s.swap(o.phi_s); r_out.swap(o.r_out);
        r_out_inv.swap(o.r_out_inv); r_search.swap(o.r_search); v_disp.swap(o.v_disp);
        time.swap(o.time); dt.swap(o.dt); acc0.swap(o.acc0);
        number.swap(o.number); rank.swap(o.rank); id_max.swap(o.id_max); id_min.swap(o.id_min);
        n_cluster.swap(o.n_cluster); flags.swap(o.flags);
        fc_rate.swap(o.fc_rate);
    }
};

// Octree + interaction lists (see oracle/orc_tree.c for the construction rules).
struct Tree {
    int n_node = 0, n_group = 0, n_level = 0;
    std::vector<int> level_begin;              // host copy, n_level+1
    DArray<uint64_t> key, key_alt;
    DArray<int32_t> perm, perm_alt, idx;
    DArray<double> bbox;                       // per-block partial mins/maxes + result
    DArray<double> origin_scale;               // origin[3], side, scale
    DArray<int32_t> node_start, node_count, node_child, node_nchild, node_level, node_parent;
    DArray<double> node_mass, node_com, node_quad, node_bmin, node_bmax, node_rs;
    DArray<int32_t> nchild_scan;
    DArray<int32_t> grp_flag, grp_scan;        // per particle
    DArray<int32_t> group_node, group_start, group_count;
    // per group: near EP particles (leaves opened by the search-radius test), far EP
    // particles (leaves opened only by the opening angle), SP nodes; each group's lists
    // are contiguous, at the per-group offsets (inside its walk workgroup's region)
    DArray<int32_t> ep_cnt, epf_cnt, sp_cnt, ep_off, epf_off, sp_off;
    DArray<int32_t> ep_list, epf_list, sp_list;
    DArray<int32_t> walk_queue;                // per-workgroup BFS queues
    DArray<int32_t> walk_use;                  // entries each walk workgroup wrote (3 per workgroup)
    DArray<int32_t> walk_err;                  // walk error bits
    DArray<long long> walk_stat;               // largest region use and total, per list
    HostBuf walk_host;
    DArray<int32_t> lvl;                       // device level ranges + overflow flag
    DArray<int32_t> lvl_part;                  // per-chunk child sums of one level
    DArray<int32_t> lvl_bnd;                   // digit bounds of one level's nodes (8 per node)
    DArray<unsigned char> scan_tmp;            // hipcub temporary storage
    int64_t n_ep_total = 0, n_epf_total = 0, n_sp_total = 0;
    int64_t n_epep_pairs = 0, n_epsp_pairs = 0;
    int queue_cap = 1 << 14;
    int n_part = 0;                            // particles of the last build
    DArray<int32_t> ng_dev;                    // the group count on the device (for the walk)
    int region_cap[3] = {4096, 4096, 4096};   // list entries per walk workgroup (EP near, EP far, SP)
    int coop_min = 1 << 16;                    // nodes of at least this many keys: workgroup-wide digit search
    // the build in progress (finished by the walk's read-back): its particle arrays, node
    // capacity and the levels launched so far
    const double *src_pos = nullptr, *src_mass = nullptr, *src_rs = nullptr;
    int src_n = 0, cap = 0, L_end = 0;
};

}  // namespace gplum
