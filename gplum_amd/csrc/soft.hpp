// soft.hpp -- EP-EP / EP-SP soft-force kernels for gfx950.
//
// One workgroup (256 threads = 4 wave64) per i-group.  Each lane holds one
// i-particle in registers; when the group has fewer than 256 particles the j-list
// is split across 2 or 4 lane sets (wave-uniform split) and the partial sums are
// combined through LDS at the end.  j-particles / super-particles are staged in LDS
// tiles of 256 as FP32 coordinates relative to the group's reference particle
// (difference formed in FP64, rounded once: gravity_kernel_epep.pikg:53,59) and
// read back as wave-wide broadcasts.  FP32 partial sums are added into FP64
// accumulators every 32 pairs per lane.  Neighbour candidates (pairs with
// |r|^2+eps^2 < 1.0201 max(r_search)^2) are compacted per wave with a ballot and a
// single atomic per wave, into a global buffer or into the group's own segment.
#pragma once
#include "common.hpp"

namespace gplum {

constexpr int SOFT_BLOCK = 256;
constexpr int SOFT_TILE = 256;

// One i-group of a batch.  Positions are taken relative to particle i_ref.
struct SoftGroup {
    int32_t i_begin, ni;      // i-particles [i_begin, i_begin+ni)
    int32_t j_begin, nj;      // near EP entries [j_begin, j_begin+nj) of the j source
    int32_t jf_begin, njf;    // far EP entries (beyond 1.1 max r_search of the group)
    int32_t s_begin, ns;      // SP entries  [s_begin, s_begin+ns) of the s source
    int32_t i_ref;            // index of the reference particle for relative coords
    int32_t pad;
};

// Packed particle record the engine keeps per particle (Morton order).
struct JRec {
    double x, y, z, m;
    double ro, rs;            // r_out, r_search
    int32_t id, rank;         // id_local (engine index) and rank (0 on one GPU)
};
static_assert(sizeof(JRec) == 56, "JRec");

// Candidate output.  Global mode (seg_off == nullptr): one counter for the launch,
// pairs[0..cap).  Segment mode: group g appends to pairs[seg_off[g] ..) through a
// counter in LDS, stores at most seg_cap[g] pairs and writes its exact count to
// seg_cnt[g] (a count above seg_cap[g] marks the group as overflowed).
struct CandBuf {
    int2 *pairs;              // (i id_local, j id_local)
    unsigned int *count;      // number of candidates found (may exceed cap)
    unsigned int cap;
    const int32_t *seg_off = nullptr, *seg_cap = nullptr;
    int32_t *seg_cnt = nullptr;
    unsigned int base = 0;    // first slot of this block's output (segment mode)
};

// ---------------------------------------------------------------- sources
// Every source returns FP64 data; the kernels derive FP32 values exactly as the
// oracle does: ro2 = (float)ro*(float)ro, rs2 = ((float)rs*(float)rs)*1.0201f.

struct SrcEPI {                       // reference EPIGrav array (C-ABI path)
    const gplum_epi *p;
    __device__ void load(int i, double &x, double &y, double &z, double &ro, double &rs,
                         int &id, int &rank) const {
        const gplum_epi &q = p[i];
        x = q.pos[0]; y = q.pos[1]; z = q.pos[2]; ro = q.r_out; rs = q.r_search;
        id = q.id_local; rank = q.myrank;
    }
};
struct SrcEPJ {                       // reference EPJGrav array (C-ABI path)
    const gplum_epj *p;
    __device__ void load(int e, double &x, double &y, double &z, double &m, double &ro,
                         double &rs, int &id, int &rank) const {
        const gplum_epj &q = p[e];
        x = q.pos[0]; y = q.pos[1]; z = q.pos[2]; m = q.mass; ro = q.r_out; rs = q.r_search;
        id = q.id_local; rank = q.myrank;
    }
};
struct SrcEPJIdx {                    // EPJGrav array by index list (MultiWalk Index)
    const gplum_epj *p;
    const int32_t *idx;
    __device__ void load(int e, double &x, double &y, double &z, double &m, double &ro,
                         double &rs, int &id, int &rank) const {
        const gplum_epj &q = p[idx[e]];
        x = q.pos[0]; y = q.pos[1]; z = q.pos[2]; m = q.mass; ro = q.r_out; rs = q.r_search;
        id = q.id_local; rank = q.myrank;
    }
};
struct SrcRecI {                      // engine i-particles: contiguous JRec
    const JRec *r;
    __device__ void load(int i, double &x, double &y, double &z, double &ro, double &rs,
                         int &id, int &rank) const {
        const JRec &q = r[i];
        x = q.x; y = q.y; z = q.z; ro = q.ro; rs = q.rs; id = q.id; rank = q.rank;
    }
};
struct SrcRecJ {                      // engine j-particles: JRec by index list
    const JRec *r;
    const int32_t *list;
    __device__ void load(int e, double &x, double &y, double &z, double &m, double &ro,
                         double &rs, int &id, int &rank) const {
        const JRec &q = r[list[e]];
        x = q.x; y = q.y; z = q.z; m = q.m; ro = q.ro; rs = q.rs; id = q.id; rank = q.rank;
    }
};
struct SrcSPQ {                       // FDPS SPJQuadrupole array
    const gplum_spj_quad *p;
    __device__ void load(int e, double &m, double &x, double &y, double &z, double q[6]) const {
        const gplum_spj_quad &s = p[e];
        m = s.mass; x = s.pos[0]; y = s.pos[1]; z = s.pos[2];
        for (int k = 0; k < 6; k++) q[k] = s.quad[k];
    }
};
struct SrcSPM {                       // FDPS SPJMonopole array
    const gplum_spj_mono *p;
    __device__ void load(int e, double &m, double &x, double &y, double &z, double q[6]) const {
        const gplum_spj_mono &s = p[e];
        m = s.mass; x = s.pos[0]; y = s.pos[1]; z = s.pos[2];
        for (int k = 0; k < 6; k++) q[k] = 0.;
    }
};
template <class SP>
struct SrcSPIdx {                     // FDPS SPJ array by index list (MultiWalk Index)
    const SP *p;
    const int32_t *idx;
    __device__ void load(int e, double &m, double &x, double &y, double &z, double q[6]) const {
        const SP &s = p[idx[e]];
        m = s.mass; x = s.pos[0]; y = s.pos[1]; z = s.pos[2];
        for (int k = 0; k < 6; k++) q[k] = quad_of(s, k);
    }
    __device__ static double quad_of(const gplum_spj_quad &s, int k) { return s.quad[k]; }
    __device__ static double quad_of(const gplum_spj_mono &, int) { return 0.; }
};
struct SrcNodeSP {                    // engine: tree-node moments by index list
    const double *mass, *com, *quad;  // com: 3 per node, quad: 6 per node
    const int32_t *list;
    __device__ void load(int e, double &m, double &x, double &y, double &z, double q[6]) const {
        const int c = list[e];
        m = mass[c]; x = com[3 * c]; y = com[3 * c + 1]; z = com[3 * c + 2];
        for (int k = 0; k < 6; k++) q[k] = quad[6 * c + k];
    }
};
struct SrcNone {
    __device__ void load(int, double &m, double &x, double &y, double &z, double q[6]) const {
        m = x = y = z = 0.;
        for (int k = 0; k < 6; k++) q[k] = 0.;
    }
    __device__ void load(int, double &x, double &y, double &z, double &m, double &ro,
                         double &rs, int &id, int &rank) const {
        x = y = z = m = ro = rs = 0.; id = rank = 0;
    }
};

// ------------------------------------------------------------------ sinks
struct SinkForce {                    // accumulate into reference ForceGrav (PIKG)
    gplum_force *f;
    __device__ void write(int i, double ax, double ay, double az, double pot, int ngb,
                          int rankf, int idmax, int idmin, bool has_ep) const {
        gplum_force &q = f[i];
        q.acc[0] += (float)ax; q.acc[1] += (float)ay; q.acc[2] += (float)az;
        q.phi += (float)pot;
        if (has_ep) {
            q.neighbor.number += ngb;
            q.neighbor.rank += rankf;
            if (idmax > q.neighbor.id_max) q.neighbor.id_max = idmax;
            if (idmin < q.neighbor.id_min) q.neighbor.id_min = idmin;
        }
    }
};
struct SinkF64 {                      // FP64 sums (+ optional Force, PIKG accumulation)
    double *acc, *phi;
    gplum_force *f;
    __device__ void write(int i, double ax, double ay, double az, double pot, int ngb,
                          int rankf, int idmax, int idmin, bool has_ep) const {
        acc[3 * i] += ax; acc[3 * i + 1] += ay; acc[3 * i + 2] += az; phi[i] += pot;
        if (f) {
            gplum_force &q = f[i];
            q.acc[0] += (float)ax; q.acc[1] += (float)ay; q.acc[2] += (float)az;
            q.phi += (float)pot;
            if (has_ep) {
                q.neighbor.number += ngb;
                q.neighbor.rank += rankf;
                if (idmax > q.neighbor.id_max) q.neighbor.id_max = idmax;
                if (idmin < q.neighbor.id_min) q.neighbor.id_min = idmin;
            }
        }
    }
};
struct SinkDiscard {                  // candidate-only passes: results already written
    __device__ void write(int, double, double, double, double, int, int, int, int, bool) const {}
};
struct SinkEngine {                   // engine particle arrays (assign, not accumulate)
    double *acc, *phi;
    int32_t *number, *rank, *id_max, *id_min;
    __device__ void write(int i, double ax, double ay, double az, double pot, int ngb,
                          int rankf, int idmax, int idmin, bool) const {
        acc[3 * i] = ax; acc[3 * i + 1] = ay; acc[3 * i + 2] = az; phi[i] = pot;
        number[i] = ngb; rank[i] = rankf; id_max[i] = idmax; id_min[i] = idmin;
    }
};

// --------------------------------------------------------- launch helpers
// NOFAR: skip the far EP list (it holds no neighbour by construction); used by the
// candidate-only pass that refills overflowed segments.
enum SoftMode : int { EP = 1, SPQ = 2, SPM = 4, EMIT = 8, NOFAR = 16 };

// tile: j-particles staged in LDS per pass (a multiple of 64 up to SOFT_TILE; the
// autotuner's choice).  The FP32 partial sums restart at tile boundaries, so the tile
// width moves results within the FP32 tolerance; neighbour flags do not change.
template <class IS, class JS, class SS, class OUT, int MODE>
void launch_soft_f32(const SoftGroup *d_groups, int ngroups, IS is, JS js, JS jf, SS ss, OUT out,
                     float eps2, CandBuf cand, hipStream_t st, int tile = SOFT_TILE);
template <class IS, class JS, class SS, class OUT, int MODE>
void launch_soft_f64(const SoftGroup *d_groups, int ngroups, IS is, JS js, JS jf, SS ss, OUT out,
                     double eps2, CandBuf cand, hipStream_t st);

}  // namespace gplum
