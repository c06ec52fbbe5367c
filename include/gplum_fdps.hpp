// gplum_fdps.hpp -- C++ host mirror of the reference's soft-force plugin interface,
// implemented over the C ABI of libgplum_gpu.so.
//
// In the reference, src/gravity_kernel.hpp:8-240 wraps the PIKG-generated functors
// CalcForceLongEPEP / CalcForceLongEPSP (names fixed by src/Makefile:70-71) in
// calcForceEPEPWithSearch / calcForceEPSP, which FDPS calls as
//   functor(const EPI_t* epi, int ni, const EPJ_t* epj, int nj, Force_t* force)
// and, for the multi-walk tree traversal, through dispatch / retrieve functions.
// The classes below keep those names, argument meanings and the reference's error
// behaviour (a fatal error aborts, as PS::Abort does); the user's EPIGrav /
// EPJGrav / SPJQuadrupole / ForceGrav types must have the reference layouts
// (individual cut-off build), which the static_asserts check.
#pragma once

#include <cstdio>
#include <cstdlib>

#include "gplum_gpu.h"

namespace gplum_fdps {

inline void check_or_abort(int rc, const char *what) {
    if (rc != GPLUM_OK) {
        std::fprintf(stderr, " ****** %s failed: %s ******\n", what, gplum_last_error());
        std::abort();
    }
}

template <class EPI, class EPJ, class FORCE>
inline void check_layouts() {
    static_assert(sizeof(EPI) == sizeof(gplum_epi), "EPI_t must have the EPIGrav layout");
    static_assert(sizeof(EPJ) == sizeof(gplum_epj), "EPJ_t must have the EPJGrav layout");
    static_assert(sizeof(FORCE) == sizeof(gplum_force), "Force_t must have the ForceGrav layout");
}

// Replacement of the PIKG functor CalcForceLongEPEP(eps2) (gravity_kernel.hpp:17-22).
struct CalcForceLongEPEP {
    float eps2;
    explicit CalcForceLongEPEP(float eps2_) : eps2(eps2_) {}
    template <class EPI, class EPJ, class FORCE>
    void operator()(const EPI *epi, const int ni, const EPJ *epj, const int nj, FORCE *force) const {
        check_layouts<EPI, EPJ, FORCE>();
        check_or_abort(gplum_calc_force_epep(reinterpret_cast<const gplum_epi *>(epi), ni,
                                             reinterpret_cast<const gplum_epj *>(epj), nj,
                                             reinterpret_cast<gplum_force *>(force), eps2),
                       "CalcForceLongEPEP");
    }
};

// Replacement of the PIKG functor CalcForceLongEPSP(eps2) (gravity_kernel.hpp:134-135),
// quadrupole super-particles (USE_QUAD).
struct CalcForceLongEPSP {
    float eps2;
    explicit CalcForceLongEPSP(float eps2_) : eps2(eps2_) {}
    template <class EPI, class SPJ, class FORCE>
    void operator()(const EPI *epi, const int ni, const SPJ *spj, const int nj, FORCE *force) const {
        static_assert(sizeof(SPJ) == sizeof(gplum_spj_quad) || sizeof(SPJ) == sizeof(gplum_spj_mono),
                      "SPJ_t must have the SPJQuadrupole or SPJMonopole layout");
        if (sizeof(SPJ) == sizeof(gplum_spj_quad))
            check_or_abort(gplum_calc_force_epsp_quad(reinterpret_cast<const gplum_epi *>(epi), ni,
                                                      reinterpret_cast<const gplum_spj_quad *>(spj), nj,
                                                      reinterpret_cast<gplum_force *>(force), eps2),
                           "CalcForceLongEPSP");
        else
            check_or_abort(gplum_calc_force_epsp_mono(reinterpret_cast<const gplum_epi *>(epi), ni,
                                                      reinterpret_cast<const gplum_spj_mono *>(spj), nj,
                                                      reinterpret_cast<gplum_force *>(force), eps2),
                           "CalcForceLongEPSP");
    }
};

// calcForceEPEPWithSearch / calcForceEPSP of src/gravity_kernel.hpp:8-240 (USE_PIKG
// branch), with the static eps2 the reference reads from FP_t::eps2.
template <class FP>
struct calcForceEPEPWithSearch {
    template <class EPI, class EPJ, class FORCE>
    void operator()(const EPI *epi, const int ni, const EPJ *epj, const int nj, FORCE *force) const {
        CalcForceLongEPEP func((float)FP::eps2);
        func(epi, ni, epj, nj, force);
    }
};
template <class FP>
struct calcForceEPSP {
    template <class EPI, class SPJ, class FORCE>
    void operator()(const EPI *epi, const int ni, const SPJ *spj, const int nj, FORCE *force) const {
        CalcForceLongEPSP func((float)FP::eps2);
        func(epi, ni, spj, nj, force);
    }
};

// Dispatch / retrieve pair for PS::TreeForForce::calcForceAllAndWriteBackMultiWalk,
// with quadrupole (USE_QUAD) or monopole super-particles.
template <class EPI, class EPJ, class SPJ>
int DispatchKernel(const int tag, const int n_walk, const EPI **epi, const int *n_epi, const EPJ **epj,
                   const int *n_epj, const SPJ **spj, const int *n_spj) {
    check_layouts<EPI, EPJ, gplum_force>();
    static_assert(sizeof(SPJ) == sizeof(gplum_spj_quad) || sizeof(SPJ) == sizeof(gplum_spj_mono),
                  "SPJ_t must have the SPJQuadrupole or SPJMonopole layout");
    const gplum_epi **e = reinterpret_cast<const gplum_epi **>(epi);
    const gplum_epj **j = reinterpret_cast<const gplum_epj **>(epj);
    if (sizeof(SPJ) == sizeof(gplum_spj_quad))
        check_or_abort(gplum_multiwalk_dispatch(tag, n_walk, e, n_epi, j, n_epj,
                                                reinterpret_cast<const gplum_spj_quad **>(spj), n_spj),
                       "DispatchKernel");
    else
        check_or_abort(gplum_multiwalk_dispatch_mono(tag, n_walk, e, n_epi, j, n_epj,
                                                     reinterpret_cast<const gplum_spj_mono **>(spj), n_spj),
                       "DispatchKernel");
    return 0;
}

// Dispatch functor of PS::TreeForForce::calcForceAllAndWriteBackMultiWalkIndex: index
// lists into the tree's whole EPJ / SPJ arrays, which are sent when send_flag is set.
template <class EPI, class EPJ, class SPJ>
int DispatchKernelIndex(const int tag, const int n_walk, const EPI **epi, const int *n_epi, const int **id_epj,
                        const int *n_epj, const int **id_spj, const int *n_spj, const EPJ *epj, const int n_epj_tot,
                        const SPJ *spj, const int n_spj_tot, const bool send_flag) {
    check_layouts<EPI, EPJ, gplum_force>();
    static_assert(sizeof(SPJ) == sizeof(gplum_spj_quad) || sizeof(SPJ) == sizeof(gplum_spj_mono),
                  "SPJ_t must have the SPJQuadrupole or SPJMonopole layout");
    static_assert(sizeof(int) == sizeof(int32_t), "index lists are S32");
    const gplum_epi **e = reinterpret_cast<const gplum_epi **>(epi);
    const int32_t **ij = reinterpret_cast<const int32_t **>(id_epj);
    const int32_t **is = reinterpret_cast<const int32_t **>(id_spj);
    const gplum_epj *j = reinterpret_cast<const gplum_epj *>(epj);
    if (sizeof(SPJ) == sizeof(gplum_spj_quad))
        check_or_abort(gplum_multiwalk_dispatch_index(tag, n_walk, e, n_epi, ij, n_epj, is, n_spj, j, n_epj_tot,
                                                      reinterpret_cast<const gplum_spj_quad *>(spj), n_spj_tot,
                                                      send_flag ? 1 : 0),
                       "DispatchKernelIndex");
    else
        check_or_abort(gplum_multiwalk_dispatch_index_mono(tag, n_walk, e, n_epi, ij, n_epj, is, n_spj, j,
                                                           n_epj_tot, reinterpret_cast<const gplum_spj_mono *>(spj),
                                                           n_spj_tot, send_flag ? 1 : 0),
                       "DispatchKernelIndex");
    return 0;
}
template <class FORCE>
int RetrieveKernel(const int tag, const int n_walk, const int *ni, FORCE **force) {
    static_assert(sizeof(FORCE) == sizeof(gplum_force), "Force_t must have the ForceGrav layout");
    check_or_abort(gplum_multiwalk_retrieve(tag, n_walk, ni, reinterpret_cast<gplum_force **>(force)),
                   "RetrieveKernel");
    return 0;
}

}  // namespace gplum_fdps
